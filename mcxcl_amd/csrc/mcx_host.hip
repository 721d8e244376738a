/*
 * mcx_host.hip - host runtime of the MI355X build: device discovery, buffer
 * staging, kernel launch, read-back and normalisation behind the reference's
 * C entry points (mcx_run_simulation / mcx_list_gpu, src/mcx_host.cpp:252-1849).
 *
 * Device memory layout (one simulation, all resident in HBM):
 *   media    uint32[nvox]          labels, x fastest, bit 31 = detector mask
 *   prop     float4[medianum]      {mua, mus, g, n}; staged into LDS per block
 *   field    double[fieldlen]      fieldlen = nvox * gates (* sources * detectors)
 *   nseed    uint32[4*nthread]     per-thread xorshift128+ seed words
 *   genergy  float[2*nthread]      escaped / launched weight per thread
 *   detrec   float[maxdet*reclen]  compacted detected-photon records
 *   detseed  uint64[2*maxdet]      launch RNG state per detected photon
 * All time gates are simulated in a single launch (288 GB of HBM removes the
 * reference's maxgate windowing, src/mcx_host.cpp:628-632, :1080).
 */
#include <cfloat>
#include <math.h>
#include <setjmp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <hip/hip_runtime.h>

#include "mcx_core.h"
#include "mcx_host.h"

/* lanes per workgroup of the bit-reproducible transport kernel */
#define MCX_LAUNCH_BLOCK ((size_t)MCX_MAX_BLOCK)
/* LDS deposit box: largest edge in voxels, the byte budget of the
   bit-reproducible build and the batched launch, and the per-workgroup LDS cap
   (gfx950: 160 KB) */
#define MCX_BOX_EDGE_MAX 48u
#define MCX_BOX_BYTES (32u * 1024u)
#define MCX_LDS_LIMIT (160u * 1024u)
/* LDS kept back from the fast build's box budget: static LDS and allocation
   granularity */
#define MCX_LDS_SLACK 1024u
/* global accumulator replicas (power of two) and their total byte cap */
#define MCX_FIELD_REPLICAS 1u
#define MCX_REPLICA_BYTES ((size_t)1 << 30)

extern "C" void* mcx_push_errjmp(void* jb);
extern "C" void mcx_pop_errjmp(void* old);
extern "C" hipError_t mcx_launch_fpmath_parity(int which, const float* x, float* y, size_t n, hipStream_t s);
extern "C" hipError_t mcx_launch_fpmath_fast(int which, const float* x, float* y, size_t n, hipStream_t s);
extern "C" hipError_t mcx_launch_rngstream(const uint32_t* seed4, unsigned int n, float* out, hipStream_t s);

/* glibc random() TYPE_3 additive-feedback generator, written from its published
   description: r[0]=seed, r[i]=16807*r[i-1] mod (2^31-1) for i<31,
   r[i]=r[i-31] for 31<=i<34, r[i]=r[i-31]+r[i-3] (mod 2^32) afterwards; the k-th
   output of rand() is r[k+344]>>1.  srand(0) behaves as srand(1).
   The state keeps the last 34 values, so a caller can draw the stream in
   pieces (one piece per repetition of a respin run). */
namespace {

struct GlibcRand {
    uint32_t ring[34];
    size_t pos;   /* index of the next value */

    explicit GlibcRand(unsigned int seed) {
        int32_t s = (int32_t)(seed == 0 ? 1 : seed);
        ring[0] = (uint32_t)s;

        for (int i = 1; i < 31; i++) {
            int32_t hi = s / 127773, lo = s % 127773;
            int32_t word = 16807 * lo - 2836 * hi;

            if (word < 0) {
                word += 2147483647;
            }

            s = word;
            ring[i] = (uint32_t)s;
        }

        for (int i = 31; i < 34; i++) {
            ring[i] = ring[i - 31];
        }

        pos = 34;

        while (pos < 344) {
            step();
        }
    }
    uint32_t step() {
        uint32_t v = ring[(pos - 31) % 34] + ring[(pos - 3) % 34];
        ring[pos % 34] = v;
        pos++;
        return v;
    }
    /* the next n outputs of rand() */
    void fill(uint32_t* out, size_t n) {
        for (size_t i = 0; i < n; i++) {
            out[i] = step() >> 1;
        }
    }
};

}  // namespace

extern "C" void mcx_host_seed_words(unsigned int seed, size_t n, unsigned int* out) {
    GlibcRand r(seed);
    r.fill(out, n);
}

extern "C" int mcx_device_count(void) {
    int n = 0;

    if (hipGetDeviceCount(&n) != hipSuccess) {
        return 0;
    }

    return n;
}

/* reference src/mcx_host.cpp:252-432 (device enumeration and autopilot sizing,
   AMD rule: 64-lane blocks, 2560 threads per compute unit, :607-609) */
extern "C" int mcx_list_gpu(Config* cfg, unsigned int* activedev, int* activedevlist, GPUInfo** info) {
    int n = mcx_device_count();
    *activedev = 0;

    if (n <= 0) {
        if (info) {
            *info = NULL;
        }

        return 0;
    }

    GPUInfo* gi = (GPUInfo*)calloc(n, sizeof(GPUInfo));

    for (int i = 0; i < n; i++) {
        hipDeviceProp_t prop;

        if (hipGetDeviceProperties(&prop, i) != hipSuccess) {
            continue;
        }

        snprintf(gi[i].name, MAX_SESSION_LENGTH, "%s (%s)", prop.name, prop.gcnArchName);
        gi[i].id = i + 1;
        gi[i].devcount = n;
        gi[i].major = prop.major;
        gi[i].minor = prop.minor;
        gi[i].globalmem = prop.totalGlobalMem;
        gi[i].constmem = prop.totalConstMem;
        gi[i].sharedmem = prop.sharedMemPerBlock;
        gi[i].regcount = prop.regsPerBlock;
        gi[i].clock = prop.clockRate;
        gi[i].sm = prop.multiProcessorCount;
        gi[i].core = prop.multiProcessorCount * 64;
        gi[i].autoblock = 64;
        gi[i].autothread = (size_t)2560 * prop.multiProcessorCount;
        gi[i].maxmpthread = prop.maxThreadsPerMultiProcessor;
        gi[i].vendor = dvAMD;
        gi[i].iscpu = 0;

        /* gpuid 0 selects the first device; deviceid[] is a '1'/'0' mask */
        int selected = (cfg && cfg->gpuid > 0) ? (cfg->gpuid == i + 1) :
                       (cfg ? (i < MAX_DEVICE && cfg->deviceid[i] == '1') : (i == 0));

        if (selected && activedevlist) {
            activedevlist[(*activedev)] = i;
        }

        if (selected) {
            (*activedev)++;
        }
    }

    if (info) {
        *info = gi;
    } else {
        free(gi);
    }

    return n;
}

/* ---------------- simulation ---------------- */
namespace {

struct DevBuf {
    void* ptr[16];
    int n;
    void add(void* p) {
        ptr[n++] = p;
    }
    void release() {
        for (int i = 0; i < n; i++) {
            if (ptr[i]) {
                (void)hipFree(ptr[i]);
            }
        }

        n = 0;
    }
};

double wall_ms() {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

}  // namespace

#define HIP_TRY(call)                                   \
    do {                                                \
        hipError_t e__ = (call);                        \
        if (e__ != hipSuccess) {                        \
            err = e__;                                  \
            errline = __LINE__;                         \
            goto fail;                                  \
        }                                               \
    } while (0)

template <typename T>
static hipError_t dalloc(DevBuf& bufs, T** p, size_t n) {
    *p = NULL;

    if (n == 0) {
        return hipSuccess;
    }

    hipError_t e = hipMalloc((void**)p, n * sizeof(T));

    if (e == hipSuccess) {
        bufs.add(*p);
    }

    return e;
}

/* features the transport kernels do not implement stop here, before any
   device work, instead of running with different semantics */
void mcx_check_supported(const Config* cfg) {
    /* packed media words (the property formats of src/mcx_core.cl:1079-1193) and
       split voxels (svmc, :1229-1348); asgn_float and mixlabel are not built */
    if (cfg->mediabyte > 4 && cfg->mediabyte != MEDIA_2LABEL_SPLIT &&
            (cfg->mediabyte < MEDIA_LABEL_HALF || cfg->mediabyte > MEDIA_AS_SHORT)) {
        MCX_ERROR(-4, "unsupported media format for the GPU kernel (label byte/short/integer, svmc, labelplus, "
                  "muamus_float, mua_float, muamus_half, asgn_byte and muamus_short are supported)");
    }

    if (cfg->mediabyte > 4 && (cfg->seed == SEED_FROM_FILE || cfg->muamus || MCX_IS_ADJOINT_TYPE(cfg->outputtype))) {
        MCX_ERROR(-4, "replay, per-voxel mua/mus arrays and adjoint outputs need a label volume");
    }

    if (cfg->mediabyte == MEDIA_LABEL_HALF && (cfg->savedetflag & 0xEu)) {
        MCX_ERROR(-4, "scattering counts, partial path lengths and momentum transfer of detected photons need a "
                  "label volume (labelplus words carry a replaced property)");
    }

    if ((cfg->srctype == MCX_SRC_PATTERN || cfg->srctype == MCX_SRC_PATTERN3D) && !cfg->srcpattern) {
        MCX_ERROR(-4, "pattern sources need the srcpattern array");
    }

    if (cfg->seed == SEED_FROM_FILE && cfg->respin != 1) {
        MCX_ERROR(-4, "respin is disabled in the replay mode");
    }

    if (cfg->minenergy >= cfg->srcpos.w) {
        MCX_ERROR(-4, "minenergy must be below the source weight (no photon could ever launch)");
    }

    if (cfg->srctype == MCX_SRC_CONE && !(cfg->srcparam1.x > 0.f)) {
        MCX_ERROR(-4, "cone source needs a positive half angle in srcparam1.x");
    }

    if (cfg->srcnum > 1 && cfg->srctype != MCX_SRC_PATTERN && cfg->srctype != MCX_SRC_PATTERN3D) {
        MCX_ERROR(-4, "photon sharing (srcnum > 1) needs a pattern or pattern3d source");
    }

    if (cfg->polmedianum && !cfg->smatrix) {
        MCX_ERROR(-4, "polarized media need their Mie tables (mcx_prep_polarized)");
    }

    /* the scattering angles are drawn by rejection against I and the linear part
       of the Stokes vector: an intensity of 0 or a degree above 1 can keep the
       draws from ever being accepted */
    if (cfg->polmedianum) {
        const MCXFloat4 s0 = cfg->srciquv;

        if (!(s0.x > 0.f) || s0.y * s0.y + s0.z * s0.z + s0.w * s0.w > s0.x * s0.x * (1.f + 1e-6f)) {
            MCX_ERROR(-4, "the source Stokes vector (srciquv) needs I > 0 and a degree of polarization of at most 1");
        }
    }

    /* the exit tally indexes the detected-photon buffer by voxel (src/mcx_utils.c:1797-1806
       grows maxdetphoton to the voxel count) */
    if (cfg->issaveref > 1 && (!cfg->issavedet || cfg->maxdetphoton < cfg->dim.x * cfg->dim.y * cfg->dim.z)) {
        MCX_ERROR(-4, "issaveref > 1 needs detected-photon output with maxdetphoton of at least the voxel count");
    }
}

/* kernel parameters of one simulation on nthread lanes in workgroups of lblock:
   geometry, time gates, record layout, the specialised-path flag, the LDS deposit
   box and the accumulator replicas.  Returns the dynamic LDS bytes per workgroup;
   *reflect receives the reference's rule for the reflection kernel variant. */
size_t mcx_kparam_setup(Config* cfg, MCXKParam* gp, size_t nthread, size_t lblock, int* reflect, size_t boxbytes) {
    const unsigned int maxgate = (unsigned int)((cfg->tend - cfg->tstart) / cfg->tstep + 0.5);
    cfg->maxgate = maxgate;
    const size_t nvox = (size_t)cfg->dim.x * cfg->dim.y * cfg->dim.z;
    /* several patterns sharing each photon keep one slot each, interleaved after the voxel */
    const unsigned int npattern = ((cfg->srctype == MCX_SRC_PATTERN || cfg->srctype == MCX_SRC_PATTERN3D) &&
                                   cfg->srcnum > 1) ? cfg->srcnum : 1;
    const unsigned int nsrcslot = (cfg->srcid < 0) ? cfg->extrasrclen + 1 : npattern;
    const unsigned int ndetslot = (cfg->seed == SEED_FROM_FILE && cfg->replaydet == -1) ? cfg->detnum : 1;
    /* complex outputs (RF replay, RF forward): the imaginary half follows the real one */
    const int isrfreplay = cfg->seed == SEED_FROM_FILE && (cfg->outputtype == otRF || cfg->outputtype == otRFmus);
    const int isrfforward = cfg->omega > 0.f && cfg->seed != SEED_FROM_FILE;
    const int isrf = isrfreplay || isrfforward;
    const size_t halflen = nvox * maxgate * nsrcslot * ndetslot;
    const size_t fieldlen = halflen * (isrf ? 2 : 1);
    const unsigned int partialdata = mcx_partialdata(cfg);
    const unsigned int reclen = mcx_detreclen(cfg);

    if (fieldlen >= 0xFFFFFFFFull || nvox >= 0x7FFFFFFFull) {
        MCX_ERROR(-4, "domain too large for 32-bit voxel indexing");
    }

    MCXKParam& g = *gp;
    memset(&g, 0, sizeof(g));
    memcpy(&g.srcpos, &cfg->srcpos, 16);
    memcpy(&g.srcdir, &cfg->srcdir, 16);
    memcpy(&g.srcparam1, &cfg->srcparam1, 16);
    memcpy(&g.srcparam2, &cfg->srcparam2, 16);
    g.maxidx_x = (float)cfg->dim.x;
    g.maxidx_y = (float)cfg->dim.y;
    g.maxidx_z = (float)cfg->dim.z;
    g.dimx = cfg->dim.x;
    g.dimy = cfg->dim.y;
    g.dimz = cfg->dim.z;
    g.dimxy = cfg->dim.x * cfg->dim.y;
    g.nvox = (uint32_t)nvox;
    g.twin0 = cfg->tstart;
    g.twin1 = cfg->tend;
    g.oneoverc0 = R_C0 * cfg->unitinmm;
    g.Rtstep = 1.f / cfg->tstep;
    g.minenergy = cfg->minenergy;
    {
        float minstep = fminf(fminf(cfg->steps.x, cfg->steps.y), cfg->steps.z);
        g.minaccumtime = minstep * R_C0 * cfg->unitinmm;
    }
    g.maxdetphoton = cfg->maxdetphoton;
    g.maxmedia = cfg->medianum - 1;
    g.medianum = cfg->medianum;
    g.detnum = cfg->detnum;
    g.voidtime = (uint32_t)cfg->voidtime;
    g.srctype = (uint32_t)cfg->srctype;
    g.maxvoidstep = (uint32_t)cfg->maxvoidstep;
    g.issaveseed = cfg->issaveseed > 0;
    g.issaveref = (uint32_t)cfg->issaveref;
    g.isspecular = cfg->isspecular > 0;
    g.maxgate = maxgate;
    /* the adjoint outputs deposit fluence per source; their products are formed
       after the run (src/mcx_core.cl:2843) */
    g.outputtype = MCX_IS_ADJOINT_TYPE(cfg->outputtype) ? (uint32_t)otFluence : (uint32_t)cfg->outputtype;
    /* detectors appended as sources (adjoint outputs, srcid -2) launch from their disks */
    g.adjsrc = ((MCX_IS_ADJOINT_TYPE(cfg->outputtype) || cfg->srcid == -2) && cfg->seed != SEED_FROM_FILE &&
                cfg->detdir && cfg->detnum && cfg->extrasrclen + 1 > cfg->detnum) ?
               cfg->extrasrclen + 1 - cfg->detnum : 0;
    g.savedetflag = cfg->savedetflag;
    g.reclen = reclen;
    g.partialdata = partialdata;
    g.w0offset = partialdata + 4;
    g.gscatter = cfg->gscatter;
    g.srcnum = cfg->srcnum;
    g.doreflect = (uint32_t)cfg->isreflect;
    g.save2pt = (uint32_t)cfg->issave2pt;
    g.savedet = (uint32_t)cfg->issavedet;
    g.nphase = cfg->nphase;
    g.extrasrclen = cfg->extrasrclen;
    g.nthread = (uint32_t)nthread;
    g.debugrng = (cfg->debuglevel & MCX_DEBUG_RNG) ? 1 : 0;
    g.mediaformat = cfg->mediabyte;
    g.seed = cfg->seed;
    g.replaydet = cfg->replaydet;
    g.srcid = cfg->srcid;
    g.fieldlen = (uint32_t)fieldlen;
    g.npattern = npattern;
    g.nangle = cfg->nangle;
    g.imoff = isrf ? (uint32_t)halflen : 0u;
    g.omega = cfg->omega;
    /* lane scratch: record columns, 2 tallies, w0, source id, one weight per
       shared pattern (srcnum), the RF replay phase (cos, sin) */
    g.pplen = g.w0offset + g.srcnum + (isrfreplay ? 2u : 0u);
    /* polarised runs: the Stokes vector after them */
    g.stokesoff = g.pplen;
    g.pplen += cfg->polmedianum ? 4u : 0u;
    /* RF forward: the complex weight at the last voxel exit */
    g.rfoff = g.pplen;
    g.pplen += isrfforward ? 2u : 0u;
    /* split voxels: the current voxel's plane {nx, ny, nz, pd} */
    g.svoff = g.pplen;
    g.pplen += (cfg->mediabyte == MEDIA_2LABEL_SPLIT) ? 4u : 0u;
    /* saved seeds: the RNG state at the photon's launch */
    g.seedoff = g.pplen;
    g.pplen += (cfg->issaveseed > 0) ? 4u : 0u;
    g.rfforward = isrfforward ? 1u : 0u;
    memcpy(g.bc, cfg->bc, 12);

    {
        unsigned int is2d = (cfg->dim.x == 1 ? 1 : (cfg->dim.y == 1 ? 2 : (cfg->dim.z == 1 ? 3 : 0)));

        if (is2d) {
            is2d = is2d * (((cfg->dim.x > 1) + (cfg->dim.y > 1) + (cfg->dim.z > 1)) == 2);
        }

        g.is2d = is2d;
    }

    /* the specialised launch/scatter/deposit path: a pencil beam starting inside a
       non-void voxel of a label volume, no focus or angular launch, no extra
       sources, no replay, no trajectory records and only forward output types */
    {
        union {
            float f;
            unsigned int u;
        } lab;
        lab.f = cfg->srcparam2.w;
        g.simple = cfg->srctype == MCX_SRC_PENCIL && cfg->srcdir.w == 0.f && (lab.u & MED_MASK) != 0 &&
                   cfg->mediabyte <= 4 && !(cfg->debuglevel & (MCX_DEBUG_MOVE | MCX_DEBUG_MOVE_ONLY)) &&
                   !(cfg->omega > 0.f) && cfg->issavedet != FILL_MAXDETPHOTON && cfg->polmedianum == 0 &&
                   cfg->extrasrclen == 0 && cfg->seed != SEED_FROM_FILE && cfg->issaveref == 0 && g.is2d == 0 &&
                   cfg->nphase <= 2 && cfg->nangle == 0 && cfg->srcpos.w > 0.f && cfg->minenergy < cfg->srcpos.w &&
                   (cfg->outputtype == otFlux || cfg->outputtype == otFluence || cfg->outputtype == otEnergy ||
                    cfg->outputtype == otL);
    }

    /* the full general path: launches and outputs the common general path leaves out */
    g.full = !g.simple &&
             (cfg->srctype == MCX_SRC_PATTERN || cfg->srctype == MCX_SRC_PATTERN3D || cfg->srctype == MCX_SRC_FOURIERX ||
              cfg->srctype == MCX_SRC_FOURIERX2D || cfg->srctype == MCX_SRC_HYPERBOLOID_GAUSSIAN || cfg->nangle ||
              cfg->outputtype == otRF || cfg->outputtype == otRFmus || cfg->outputtype == otWLTOF ||
              cfg->outputtype == otWPTOF || cfg->mediabyte > 4 || (cfg->debuglevel & (MCX_DEBUG_MOVE | MCX_DEBUG_MOVE_ONLY)) ||
              cfg->polmedianum > 0 ||
              (cfg->omega > 0.f && cfg->seed != SEED_FROM_FILE));

    /* trajectory records (src/mcx_core.cl:929-944) */
    g.savetraj = (cfg->debuglevel & (MCX_DEBUG_MOVE | MCX_DEBUG_MOVE_ONLY)) ? 1u : 0u;
    g.moveonly = (cfg->debuglevel & MCX_DEBUG_MOVE_ONLY) ? 1u : 0u;
    g.maxjumpdebug = cfg->maxjumpdebug;
    /* polarised media (src/mcx_core.cl:2454-2565) and Stokes trajectory records */
    g.maxpolmedia = cfg->polmedianum;
    g.istrajstokes = cfg->istrajstokes ? 1u : 0u;
    g.trajlen = MCX_DEBUG_REC_LEN + (g.istrajstokes << 2);
    g.s0 = make_float4(cfg->srciquv.x, cfg->srciquv.y, cfg->srciquv.z, cfg->srciquv.w);
    g.svmc = (cfg->mediabyte == MEDIA_2LABEL_SPLIT) ? 1u : 0u;

    /* reflection block rule (src/mcx_host.cpp:945-956) */
    {
        char bcs[13] = {0}, allabsorb[7] = {bcAbsorb, bcAbsorb, bcAbsorb, bcAbsorb, bcAbsorb, bcAbsorb, 0};
        memcpy(bcs, cfg->bc, 12);
        *reflect = cfg->isreflect || (strcmp(bcs, allabsorb) && bcs[0] != 0);

        for (int i = 0; i < 6; i++) {
            if (cfg->bc[i] == bcReflect || cfg->bc[i] == bcMirror) {
                *reflect = 1;
            }
        }
    }

    size_t pplen = g.pplen;
    size_t shmem = cfg->medianum * sizeof(float4) + ((cfg->nphase + 3) & ~3u) * sizeof(float) +
                   lblock * pplen * sizeof(float);

    if (shmem + MCX_LDS_SLACK > MCX_LDS_LIMIT) {
        MCX_ERROR(-4, "too many media/phase samples or record columns for the LDS tables");
    }

    /* deposit privatisation: most fluence deposits, and every heavily contended
       cache line, fall in the first gates around the source.  A box centred on the
       source voxel (clamped to the grid) is accumulated per workgroup in LDS and
       added to the global grid once at the end.  Cells are float64 in the
       bit-reproducible build and float32 in the fast build (the reference
       accumulates the whole grid in float32); the budget (boxbytes, in bytes)
       buys up to three leading gates of the largest box that fits. */
    {
        const size_t elem = (cfg->mathmode == 1) ? sizeof(double) : sizeof(float);
        size_t budget = boxbytes;
        const char* env = getenv("MCXHIP_BOX_BYTES");

        if (env) {
            budget = (size_t)strtoul(env, NULL, 10);
        }

        if (g.debugrng || !cfg->issave2pt || budget < elem || npattern > 1) {
            budget = 0;   /* shared patterns deposit straight to their interleaved slots */
        }

        /* RF forward: the box holds the imaginary half of its gates too (a pattern
           source's RF field has none) */
        const int boxim = isrfforward && cfg->srctype != MCX_SRC_PATTERN && cfg->srctype != MCX_SRC_PATTERN3D;
        const size_t cells = budget / elem / (boxim ? 2 : 1);
        /* leading gates held: three when the budget buys three gates of 20^3, two for
           two of 16^3, else one */
        size_t ng = (cells >= 3u * 20u * 20u * 20u && maxgate >= 3) ? 3 :
                    ((cells >= 2u * 16u * 16u * 16u && maxgate >= 2) ? 2 : 1);
        const unsigned int dims[3] = {cfg->dim.x, cfg->dim.y, cfg->dim.z};
        float spos[3] = {cfg->srcpos.x, cfg->srcpos.y, cfg->srcpos.z};
        const float sdir[3] = {cfg->srcdir.x, cfg->srcdir.y, cfg->srcdir.z};
        /* a source on a face of the grid deposits into the half space in front of
           it: the box is flattened along that face's normal, depth 3/4 of its width
           (counting the deposits per voxel and gate of a pencil beam into the 60^3
           sphere-in-cube volume, this shape with three gates holds 45.2% of them
           against 43.7% for a cube with two) */
        int depth = -1;

        for (int i = 0; i < 3; i++) {
            const int c = (int)floorf(spos[i]);

            if ((c <= 0 || c >= (int)dims[i] - 1) && (depth < 0 || fabsf(sdir[i]) > fabsf(sdir[depth]))) {
                depth = i;
            }
        }

        unsigned int lat = 1;

        while (lat < MCX_BOX_EDGE_MAX) {
            const size_t l = lat + 1, d = (depth >= 0) ? (3 * l) / 4 : l;

            if (l * l * (d ? d : 1) * ng > cells) {
                break;
            }

            lat++;
        }

        unsigned int edge[3];
        unsigned int org[3];

        for (int i = 0; i < 3; i++) {
            unsigned int e = (i == depth) ? (3 * lat) / 4 : lat;
            e = e ? e : 1;
            edge[i] = dims[i] < e ? dims[i] : e;
            int c = (int)floorf(spos[i]) - (int)(edge[i] / 2);
            int hi = (int)(dims[i] - edge[i]);
            org[i] = (unsigned int)(c < 0 ? 0 : (c > hi ? hi : c));
        }

        size_t pergate = (size_t)edge[0] * edge[1] * edge[2];
        const size_t ngmax = ng;
        ng = (pergate && cells >= pergate) ? cells / pergate : 0;
        ng = ng < ngmax ? ng : ngmax;
        ng = ng < maxgate ? ng : maxgate;

        if (ng > 0 && shmem + 8 + ng * pergate * elem * (boxim ? 2 : 1) <= MCX_LDS_LIMIT) {
            g.box_x0 = org[0];
            g.box_y0 = org[1];
            g.box_z0 = org[2];
            g.box_nx = edge[0];
            g.box_ny = edge[1];
            g.box_nz = edge[2];
            g.box_ng = (uint32_t)ng;
            g.box_cells = (uint32_t)(ng * pergate * (boxim ? 2 : 1));
            g.box_imgate = boxim ? (uint32_t)(halflen / nvox) : 0u;
            /* the box follows the 16-byte aligned float tables */
            shmem += (size_t)g.box_cells * elem + 8;
        }
    }

    /* replicas of the global accumulator spread the remaining contended lines */
    {
        size_t nr = MCX_FIELD_REPLICAS;
        const char* env = getenv("MCXHIP_REPLICAS");

        if (env) {
            nr = (size_t)strtoul(env, NULL, 10);
        }

        if (g.debugrng) {
            nr = 1;
        }

        while (nr > 1 && (nr & (nr - 1))) {
            nr &= nr - 1;
        }

        while (nr > 1 && fieldlen * nr * sizeof(double) > MCX_REPLICA_BYTES) {
            nr >>= 1;
        }

        g.nrep = nr ? (uint32_t)nr : 1u;
    }

    return shmem;
}

/* pixels of a pattern source: param1.w x param2.w (pattern), param1.x x
   param1.y x param1.z (pattern3d) */
static size_t mcx_pattern_pixels(const Config* cfg) {
    if (cfg->srctype == MCX_SRC_PATTERN3D) {
        return (size_t)((int)cfg->srcparam1.x * (int)cfg->srcparam1.y * (int)cfg->srcparam1.z);
    }

    return (size_t)((int)cfg->srcparam1.w * (int)cfg->srcparam2.w);
}

/* normalisation factor of the float output (src/mcx_host.cpp:1382-1465, after
   the energy tallies are known): flux = unitinmm / (E_launched Vvox tstep),
   fluence = flux x tstep, energy and path length = 1 / E_launched, replay
   outputs = unitinmm / sum of the replay weights; x (number of sources) when
   every source is simulated into its own slice.  0 when replay output is
   scaled per detector afterwards. */
static float mcx_output_scale(const Config* cfg) {
    const float Vvox = cfg->steps.x * cfg->steps.y * cfg->steps.z;
    float scale = 1.f;

    if (cfg->outputtype == otFlux || cfg->outputtype == otFluence || MCX_IS_ADJOINT_TYPE(cfg->outputtype) ||
            (cfg->omega > 0.f && cfg->seed != SEED_FROM_FILE)) {
        scale = cfg->unitinmm / (float)(cfg->energytot * Vvox * cfg->tstep);

        if (cfg->outputtype == otFluence || MCX_IS_ADJOINT_TYPE(cfg->outputtype)) {
            scale *= cfg->tstep;
        }
    } else if (cfg->outputtype == otEnergy || cfg->outputtype == otL) {
        scale = 1.f / (float)cfg->energytot;
    } else if (cfg->outputtype == otJacobian || cfg->outputtype == otWP || cfg->outputtype == otDCS ||
               cfg->outputtype == otRF || cfg->outputtype == otRFmus || cfg->outputtype == otWLTOF ||
               cfg->outputtype == otWPTOF) {
        if (cfg->seed == SEED_FROM_FILE && cfg->replaydet == -1) {
            return 0.f;   /* per-detector scaling after the read-back */
        }

        float sw = 0.f;

        for (size_t i = 0; i < cfg->nphoton; i++) {
            sw += cfg->replay.weight[i];
        }

        scale = sw > 0.f ? cfg->unitinmm / sw : 0.f;
    }

    if (cfg->extrasrclen && cfg->srcid < 0) {
        scale *= (float)(cfg->extrasrclen + 1);
    }

    return scale;
}

/* Launch plan of one run on device dev (ncu compute units) with *nthread_io
   lanes of the reference's thread map: the kernel parameters, the workgroup
   size and count, the dynamic LDS bytes and the reflection variant.  Fast mode
   (outside replay) turns the thread map into one resident wave of workgroups
   that hand photons out at run time and updates *nthread_io to its lanes. */
static void mcx_plan_run(Config* cfg, int dev, int ncu, size_t* nthread_io, MCXKParam* gp, size_t* lblock_out,
                         size_t* nlaunch_out, size_t* shmem_out, int* reflect_out) {
    const char* envblock = getenv("MCXHIP_BLOCK");
    const char* envdyn = getenv("MCXHIP_DYNAMIC");
    const int fast = cfg->mathmode == 0 && !(cfg->debuglevel & MCX_DEBUG_RNG);
    /* fast mode hands photons out at run time (below) unless replaying, where
       each lane must rerun the photons of its recorded seeds, or recording
       trajectories, whose photon ids follow the fixed map */
    const int dynamic = fast && cfg->seed != SEED_FROM_FILE && !(cfg->debuglevel & (MCX_DEBUG_MOVE | MCX_DEBUG_MOVE_ONLY)) &&
                        !(envdyn && envdyn[0] == '0');
    size_t lblock = MCX_LAUNCH_BLOCK;
    MCXKParam& g = *gp;
    int reflect = 0;
    size_t nthread = *nthread_io;
    size_t shmem = mcx_kparam_setup(cfg, &g, nthread, lblock, &reflect, fast ? 0 : MCX_BOX_BYTES);
    {
        const size_t bound = fast ? (size_t)MCX_FAST_BOUND_FULL(cfg->issavedet != 0, reflect != 0, g.simple != 0,
                                 g.full != 0) : MCX_MAX_BLOCK;

        if (envblock) {
            lblock = (size_t)strtoul(envblock, NULL, 10);
        } else if (dynamic) {
            /* the wide workgroup holds the same lanes per CU as four 256-lane ones,
               so its lane scratch costs no more LDS per CU; it only has to fit in
               one workgroup's LDS (many media with full records may not) */
            const size_t lane = (size_t)g.pplen * sizeof(float);
            lblock = (shmem + (bound - MCX_LAUNCH_BLOCK) * lane + MCX_LDS_SLACK <= MCX_LDS_LIMIT) ? bound
                     : MCX_LAUNCH_BLOCK;
        }

        if (lblock < 64 || lblock > bound || (lblock & 63)) {
            MCX_ERROR(-4, "MCXHIP_BLOCK must be a multiple of 64 within the kernel's launch bounds");
        }
    }

    /* fast mode: the deposit box takes the LDS a CU has left for each of its
       resident workgroups, given the occupancy the kernel's registers allow */
    if (fast) {
        int ldscu = 0, ldsblock = 0, nwg = 0;

        if (hipSetDevice(dev) != hipSuccess) {
            MCX_ERROR(-99, "cannot select the HIP device");
        }

        shmem = mcx_kparam_setup(cfg, &g, nthread, lblock, &reflect, 0);

        if (hipDeviceGetAttribute(&ldscu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) == hipSuccess &&
                hipDeviceGetAttribute(&ldsblock, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess) {
            nwg = mcx_fast_occupancy((int)lblock, shmem, cfg->issavedet != 0, reflect, g.full ? 2 : (int)g.simple);

            if (nwg <= 0 && dynamic && !envblock && lblock > MCX_MAX_BLOCK) {
                /* no CU holds the wide workgroup: fall back to 256 lanes */
                lblock = MCX_MAX_BLOCK;
                shmem = mcx_kparam_setup(cfg, &g, nthread, lblock, &reflect, 0);
                nwg = mcx_fast_occupancy((int)lblock, shmem, cfg->issavedet != 0, reflect, g.full ? 2 : (int)g.simple);
            }
        }

        size_t budget = 0;

        if (nwg > 0 && ldscu > 0 && ldsblock > 0) {
            size_t per = (size_t)ldscu / (size_t)nwg;
            per = per < (size_t)ldsblock ? per : (size_t)ldsblock;
            per = per < MCX_LDS_LIMIT ? per : MCX_LDS_LIMIT;
            budget = per > shmem + MCX_LDS_SLACK ? per - shmem - MCX_LDS_SLACK : 0;
        }

        shmem = mcx_kparam_setup(cfg, &g, nthread, lblock, &reflect, budget);
    }

    size_t nlaunch = (nthread + lblock - 1) / lblock;
    /* respin: |respin| repetitions of nphoton/|respin| photons each, every one on
       fresh seeds drawn further down the same rand() stream
       (src/mcx_host.cpp:1089-1104, :1319-1332) */
    const unsigned int respin = (unsigned int)(cfg->respin < 0 ? -cfg->respin : (cfg->respin ? cfg->respin : 1));
    const size_t passphoton = cfg->nphoton / respin;

    /* photon split across the persistent threads (src/mcx_host.cpp:1101-1102).
       A lane counts its photons in a float (as the reference's f.w), exact
       below 2^24. */
    if (passphoton / nthread >= ((size_t)1 << 24)) {
        MCX_ERROR(-4, "too many photons per thread (more than 2^24): use more threads or respin");
    }

    g.threadphoton = (uint32_t)(passphoton / nthread);
    g.oddphoton = (int32_t)(passphoton - (size_t)g.threadphoton * nthread);

    /* fast mode: one resident wave of workgroups, each drawing its share of photons
       one at a time, so lanes stay busy until their workgroup's share is done
       instead of each running a fixed count (per-lane work varies by +-16% at
       150 photons per lane and a wave waits for its slowest lane).  Bit-parity
       mode keeps the reference's fixed photon-per-thread map. */
    {
        int nb = (dynamic && !g.debugrng) ?
                 mcx_fast_occupancy((int)lblock, shmem, cfg->issavedet != 0, reflect, g.full ? 2 : (int)g.simple) : 0;

        if (nb > 0 && ncu > 0) {
            nlaunch = (size_t)nb * (size_t)ncu;
            /* MCXHIP_MAXWG caps the workgroup count (tests use it to give each
               workgroup, and so each deposit box, many photons) */
            const char* envmaxwg = getenv("MCXHIP_MAXWG");

            if (envmaxwg && strtoul(envmaxwg, NULL, 10) > 0 && strtoul(envmaxwg, NULL, 10) < nlaunch) {
                nlaunch = (size_t)strtoul(envmaxwg, NULL, 10);
            }

            /* a workgroup's counter runs from 0 to its share plus one overshooting
               draw per lane, in 32 bits */
            if (passphoton / nlaunch + 1 + lblock >= (size_t)0xFFFFFFFFu) {
                MCX_ERROR(-4, "too many photons per workgroup (2^32): use respin to split the run");
            }

            nthread = nlaunch * lblock;
            g.nthread = (uint32_t)nthread;
            g.dynamic = 1;
            g.threadphoton = (uint32_t)(passphoton / nlaunch);
            g.oddphoton = (int32_t)(passphoton - (size_t)g.threadphoton * nlaunch);
        }
    }

    if (getenv("MCXHIP_VERBOSE")) {
        fprintf(stderr, "mcxhip: %zu lanes in %zu workgroups of %zu (%s), LDS %zu B, box %ux%ux%u x %u gates at "
                "(%u,%u,%u), %u replicas, simple %u\n", nthread, nlaunch, lblock,
                g.dynamic ? "dynamic photons" : "fixed photons per lane", shmem, g.box_nx, g.box_ny, g.box_nz,
                g.box_ng, g.box_x0, g.box_y0, g.box_z0, g.nrep, g.simple);
    }

    *nthread_io = nthread;
    *lblock_out = lblock;
    *nlaunch_out = nlaunch;
    *shmem_out = shmem;
    *reflect_out = reflect;
}

extern "C" void mcx_run_simulation(Config* cfg, float* fluence, float* totalenergy) {
    (void)fluence;
    (void)totalenergy;
    hipError_t err = hipSuccess;
    int errline = 0;
    DevBuf bufs;
    bufs.n = 0;
    hipStream_t stream = NULL;
    hipEvent_t ev0 = NULL, ev1 = NULL, ev2 = NULL;
    uint32_t* hseed = NULL;
    float* hgenergy = NULL;
    float* hfield = NULL;
    uint32_t hdetcount = 0;
    char msg[512];
    GlibcRand seedgen(cfg->seed > 0 ? (unsigned int)cfg->seed : (unsigned int)time(0));

    mcx_check_supported(cfg);

    unsigned int activedev = 0;
    int devlist[MAX_DEVICE];
    GPUInfo* gpu = NULL;
    int ndev = mcx_list_gpu(cfg, &activedev, devlist, &gpu);

    if (ndev == 0 || activedev == 0) {
        free(gpu);
        MCX_ERROR(-99, "No HIP device found (or the specified GPU does not exist)");
    }

    const int dev = devlist[0];

    /* thread map: reference autopilot for AMD devices, or the user's nthread/nblocksize */
    size_t nblocksize = cfg->autopilot ? gpu[dev].autoblock : cfg->nblocksize;
    size_t nthread = cfg->autopilot ? gpu[dev].autothread : cfg->nthread;
    const int ncu = gpu[dev].sm;

    if (nblocksize == 0) {
        free(gpu);
        MCX_ERROR(-4, "block size must be positive");
    }

    /* the reference rounds the thread count down to whole blocks; the photon map
       depends only on the thread count, so the transport kernel itself is launched
       in workgroups of its own size (lblock) with the tail lanes idle */
    nthread = (nthread / nblocksize) * nblocksize;

    if (nthread == 0) {
        free(gpu);
        MCX_ERROR(-4, "thread number must be at least one block");
    }

    free(gpu);
    gpu = NULL;

    size_t lblock = 0, nlaunch = 0, shmem = 0;
    int reflect = 0;
    MCXKParam g;
    mcx_plan_run(cfg, dev, ncu, &nthread, &g, &lblock, &nlaunch, &shmem, &reflect);
    const unsigned int maxgate = g.maxgate;
    const size_t nvox = g.nvox;
    const size_t fieldlen = g.fieldlen;
    const unsigned int reclen = g.reclen;
    const unsigned int respin = (unsigned int)(cfg->respin < 0 ? -cfg->respin : (cfg->respin ? cfg->respin : 1));

    if (cfg->seed == SEED_FROM_FILE &&
            (cfg->nphoton == 0 || !cfg->replay.seed || !cfg->replay.weight || !cfg->replay.tof || !cfg->replay.detid)) {
        MCX_ERROR(-6, "replay mode needs seeds, weights, times of flight and detector ids");
    }

    cfg->detectedcount = 0;
    cfg->his.detected = 0;
    cfg->his.respin = cfg->respin;
    cfg->his.colcount = reclen;
    cfg->energytot = 0;
    cfg->energyesc = 0;
    cfg->runtime = 0;

    double t0 = 0, t1 = 0;
    float kernelms = 0.f;
    uint32_t* dmedia = NULL;
    float4* dprop = NULL;
    float4* dsmatrix = NULL;
    float* dpolcdf = NULL;
    float* hpolcdf = NULL;
    float4* ddetpos = NULL;
    float* dinvcdf = NULL;
    float* dangle = NULL;
    float* dpattern = NULL;
    float4* dsrcdata = NULL;
    float* dreplayw = NULL;
    float* dreplaytof = NULL;
    int* dreplaydetid = NULL;
    uint64_t* dreplayseed = NULL;
    uint32_t* dnseed = NULL;
    double* dfield = NULL;
    float* dgenergy = NULL;
    float* ddetrec = NULL;
    uint64_t* ddetseed = NULL;
    float* ddetw = NULL;
    float2* dmuamus = NULL;
    uint32_t* ddetcount = NULL;
    float* ddebug = NULL;
    uint32_t* djump = NULL;
    float* dout = NULL;
    float* draw32 = NULL;
    float* djacob = NULL;
    float* dopscale = NULL;
    /* float output wanted; adjoint outputs wanted, their length and opscale use */
    int wantfloat = 0, wantadj = 0, adjdual = 0, adjop = 0, adjmusp = 0;
    size_t adjlen = 0;
    size_t nrep = 0;

    HIP_TRY(hipSetDevice(dev));
    HIP_TRY(hipStreamCreate(&stream));
    HIP_TRY(hipEventCreate(&ev0));
    HIP_TRY(hipEventCreate(&ev1));
    HIP_TRY(hipEventCreate(&ev2));

    /* split voxels: the label words, then one plane word per voxel */
    HIP_TRY(dalloc(bufs, &dmedia, nvox * (cfg->mediabyte == MEDIA_2LABEL_SPLIT ? 2 : 1)));
    HIP_TRY(hipMemcpyAsync(dmedia, cfg->vol, nvox * (cfg->mediabyte == MEDIA_2LABEL_SPLIT ? 2 : 1) * sizeof(uint32_t),
                           hipMemcpyHostToDevice, stream));
    HIP_TRY(dalloc(bufs, &dprop, cfg->medianum));

    if (cfg->polmedianum) {
        HIP_TRY(dalloc(bufs, &dsmatrix, (size_t)cfg->polmedianum * MCX_NANGLES));
        HIP_TRY(hipMemcpyAsync(dsmatrix, cfg->smatrix, (size_t)cfg->polmedianum * MCX_NANGLES * sizeof(float4),
                               hipMemcpyHostToDevice, stream));

        /* the fast build's angle sampler: theta's marginal S11(bin) sin(theta) as a CDF over
           the table's bins (bin b = floor(theta 181 (1/pi - EPS)), uniform in cos(theta)
           inside a bin) */
        const size_t ne = MCX_NANGLES + 1;
        hpolcdf = (float*)malloc(ne * (1 + cfg->polmedianum) * sizeof(float));
        const double width = 1.0 / (MCX_NANGLES * (1.0 / M_PI - (double)FLT_EPSILON));

        for (size_t b = 0; b < ne; b++) {
            hpolcdf[b] = (float)cos(fmin(b * width, M_PI));
        }

        for (unsigned int m = 0; m < cfg->polmedianum; m++) {
            float* cdf = hpolcdf + ne * (1 + m);
            double acc = 0.0;
            cdf[0] = 0.f;

            for (size_t b = 0; b < MCX_NANGLES; b++) {
                acc += (double)cfg->smatrix[m * MCX_NANGLES + b].x * ((double)hpolcdf[b] - (double)hpolcdf[b + 1]);
                cdf[b + 1] = (float)acc;
            }

            for (size_t b = 1; b < ne; b++) {
                cdf[b] = (float)(cdf[b] / acc);
            }

            cdf[MCX_NANGLES] = 1.f;
        }

        HIP_TRY(dalloc(bufs, &dpolcdf, ne * (1 + cfg->polmedianum)));
        HIP_TRY(hipMemcpyAsync(dpolcdf, hpolcdf, ne * (1 + cfg->polmedianum) * sizeof(float), hipMemcpyHostToDevice,
                               stream));
    }
    HIP_TRY(hipMemcpyAsync(dprop, cfg->prop, cfg->medianum * sizeof(float4), hipMemcpyHostToDevice, stream));

    if (cfg->detnum) {
        HIP_TRY(dalloc(bufs, &ddetpos, cfg->detnum));
        HIP_TRY(hipMemcpyAsync(ddetpos, cfg->detpos, cfg->detnum * sizeof(float4), hipMemcpyHostToDevice, stream));
    }

    if (cfg->nphase) {
        HIP_TRY(dalloc(bufs, &dinvcdf, cfg->nphase));
        HIP_TRY(hipMemcpyAsync(dinvcdf, cfg->invcdf, cfg->nphase * sizeof(float), hipMemcpyHostToDevice, stream));
    }

    if (cfg->nangle) {
        HIP_TRY(dalloc(bufs, &dangle, cfg->nangle));
        HIP_TRY(hipMemcpyAsync(dangle, cfg->angleinvcdf, cfg->nangle * sizeof(float), hipMemcpyHostToDevice, stream));
    }

    if ((cfg->srctype == MCX_SRC_PATTERN || cfg->srctype == MCX_SRC_PATTERN3D) && cfg->srcpattern) {
        const size_t npat = mcx_pattern_pixels(cfg) * g.npattern;
        HIP_TRY(dalloc(bufs, &dpattern, npat));
        HIP_TRY(hipMemcpyAsync(dpattern, cfg->srcpattern, npat * sizeof(float), hipMemcpyHostToDevice, stream));
    }

    if (cfg->extrasrclen) {
        HIP_TRY(dalloc(bufs, &dsrcdata, (size_t)cfg->extrasrclen * 4));
        HIP_TRY(hipMemcpyAsync(dsrcdata, cfg->srcdata, cfg->extrasrclen * sizeof(ExtraSrc), hipMemcpyHostToDevice, stream));
    }

    if (cfg->seed == SEED_FROM_FILE) {
        nrep = cfg->nphoton;

        HIP_TRY(dalloc(bufs, &dreplayw, nrep));
        HIP_TRY(dalloc(bufs, &dreplaytof, nrep));
        HIP_TRY(dalloc(bufs, &dreplaydetid, nrep));
        HIP_TRY(dalloc(bufs, &dreplayseed, nrep * 2));
        HIP_TRY(hipMemcpyAsync(dreplayw, cfg->replay.weight, nrep * sizeof(float), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(dreplaytof, cfg->replay.tof, nrep * sizeof(float), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(dreplaydetid, cfg->replay.detid, nrep * sizeof(int), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(dreplayseed, cfg->replay.seed, nrep * 16, hipMemcpyHostToDevice, stream));
    }

    /* seeds: glibc srand(seed) then one rand() per 32-bit word (src/mcx_host.cpp:696-700, :759-768) */
    hseed = (uint32_t*)malloc(nthread * 4 * sizeof(uint32_t));

    if (!hseed) {
        HIP_TRY(hipErrorOutOfMemory);
    }

    seedgen.fill(hseed, nthread * 4);
    HIP_TRY(dalloc(bufs, &dnseed, nthread * 4));
    HIP_TRY(hipMemcpyAsync(dnseed, hseed, nthread * 4 * sizeof(uint32_t), hipMemcpyHostToDevice, stream));

    HIP_TRY(dalloc(bufs, &dfield, fieldlen * g.nrep));
    HIP_TRY(hipMemsetAsync(dfield, 0, fieldlen * g.nrep * sizeof(double), stream));
    HIP_TRY(dalloc(bufs, &dgenergy, nthread * 2));
    HIP_TRY(hipMemsetAsync(dgenergy, 0, nthread * 2 * sizeof(float), stream));
    HIP_TRY(dalloc(bufs, &ddetcount, 1));
    HIP_TRY(hipMemsetAsync(ddetcount, 0, sizeof(uint32_t), stream));

    if (g.savetraj) {
        HIP_TRY(dalloc(bufs, &ddebug, (size_t)cfg->maxjumpdebug * g.trajlen));
        HIP_TRY(dalloc(bufs, &djump, 1));
        HIP_TRY(hipMemsetAsync(djump, 0, sizeof(uint32_t), stream));
    }

    if (cfg->issavedet) {
        HIP_TRY(dalloc(bufs, &ddetrec, (size_t)cfg->maxdetphoton * (reclen ? reclen : 1)));

        if (cfg->issaveref > 1) {
            /* the exit tally accumulates into the records, one per voxel */
            HIP_TRY(hipMemsetAsync(ddetrec, 0, (size_t)cfg->maxdetphoton * reclen * sizeof(float), stream));
        }

        if (cfg->issaveseed) {
            HIP_TRY(dalloc(bufs, &ddetseed, (size_t)cfg->maxdetphoton * 2));
        }

        if (cfg->exportdetweight) {
            HIP_TRY(dalloc(bufs, &ddetw, (size_t)cfg->maxdetphoton));
        }
    }

    if (cfg->muamus) {
        /* per-voxel mua/mus, converted from 1/mm to 1/grid unit as the label table is */
        float2* hm = (float2*)malloc(nvox * sizeof(float2));

        if (!hm) {
            HIP_TRY(hipErrorOutOfMemory);
        }

        for (size_t i = 0; i < nvox; i++) {
            hm[i].x = cfg->muamus[2 * i] * cfg->unitinmm;
            hm[i].y = cfg->muamus[2 * i + 1] * cfg->unitinmm;
        }

        hipError_t e = dalloc(bufs, &dmuamus, nvox);

        if (e == hipSuccess) {
            e = hipMemcpy(dmuamus, hm, nvox * sizeof(float2), hipMemcpyHostToDevice);
        }

        free(hm);
        HIP_TRY(e);
    }

    g.media = dmedia;
    g.prop = dprop;
    g.smatrix = dsmatrix;
    g.polcdf = dpolcdf;
    g.detpos = ddetpos;
    g.invcdf = dinvcdf;
    g.angleinvcdf = dangle;
    g.srcpattern = dpattern;
    g.srcdata = dsrcdata;
    g.replayweight = dreplayw;
    g.replaytof = dreplaytof;
    g.replaydetid = dreplaydetid;
    g.replayseed = dreplayseed;
    g.nseed = dnseed;
    g.field = dfield;
    g.genergy = dgenergy;
    g.detrec = ddetrec;
    g.detseed = ddetseed;
    g.detcount = ddetcount;
    g.detweight = ddetw;
    g.debugdata = ddebug;
    g.jumpdebug = djump;
    g.muamus = dmuamus;

    /* the post-processing buffers are allocated before the timed region: a
       caller that hands in only the float64 or float32 raw grid (exportfield
       NULL) gets no normalised float32 copy, so large replay grids are read
       back once */
    if (cfg->issave2pt) {
        wantfloat = cfg->exportfield != NULL || (cfg->exportfield64 == NULL && cfg->exportfield32 == NULL) ||
                    cfg->parentid == mpStandalone;
        wantadj = wantfloat && MCX_IS_ADJOINT_TYPE(cfg->outputtype) && g.adjsrc;

        if (wantfloat) {
            HIP_TRY(dalloc(bufs, &dout, fieldlen));
        }

        if (wantadj) {
            adjlen = nvox * g.adjsrc * cfg->detnum;
            adjdual = MCX_IS_DUAL_ADJOINT_TYPE(cfg->outputtype);
            adjmusp = cfg->outputtype == otAdjointMusp || cfg->outputtype == otAdjointMuaMusp;
            adjop = adjmusp || cfg->outputtype == otAdjointMus;
            /* RF forward: a real and an imaginary block per output (src/mcx_host.cpp:1474-1476) */
            HIP_TRY(dalloc(bufs, &djacob, adjlen * (adjdual ? 2 : 1) * (g.rfforward ? 2 : 1)));

            if (adjop) {
                HIP_TRY(dalloc(bufs, &dopscale, 256));
            }
        }

        if (cfg->exportfield32) {
            HIP_TRY(dalloc(bufs, &draw32, fieldlen));
        }
    }

    HIP_TRY(hipStreamSynchronize(stream));
    t0 = wall_ms();

    /* one launch per repetition: the grid, the energy tallies and the detected
       records accumulate on the device; each repetition after the first gets the
       next nthread*4 rand() words as seeds (src/mcx_host.cpp:1319-1332) */
    for (unsigned int pass = 0; pass < respin; pass++) {
        float passms = 0.f;

        if (pass > 0) {
            seedgen.fill(hseed, nthread * 4);
            HIP_TRY(hipMemcpyAsync(dnseed, hseed, nthread * 4 * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
            HIP_TRY(hipStreamSynchronize(stream));
        }

        HIP_TRY(hipEventRecord(ev0, stream));

        if (g.debugrng) {
            HIP_TRY(mcx_launch_rngdump(&g, (int)nlaunch, (int)lblock, stream));
        } else if (cfg->mathmode == 1) {
            HIP_TRY(mcx_launch_parity(&g, (int)nlaunch, (int)lblock, shmem, stream, cfg->issavedet != 0,
                                      reflect));
        } else {
            HIP_TRY(mcx_launch_fast(&g, (int)nlaunch, (int)lblock, shmem, stream, cfg->issavedet != 0,
                                    reflect));
        }

        HIP_TRY(hipEventRecord(ev2, stream));
        HIP_TRY(hipEventSynchronize(ev2));
        HIP_TRY(hipEventElapsedTime(&passms, ev0, ev2));
        kernelms += passms;
    }

    t1 = wall_ms();
    cfg->kernelms = kernelms;
    cfg->jobms = kernelms;
    cfg->runtimems = t1 - t0;
    cfg->runtime = (unsigned int)(t1 - t0 + 0.5);

    /* energy tallies, summed in thread order in double (src/mcx_host.cpp:1299-1306) */
    hgenergy = (float*)malloc(nthread * 2 * sizeof(float));
    HIP_TRY(hipMemcpy(hgenergy, dgenergy, nthread * 2 * sizeof(float), hipMemcpyDeviceToHost));

    for (size_t i = 0; i < nthread; i++) {
        cfg->energyesc += hgenergy[i * 2];
        cfg->energytot += hgenergy[i * 2 + 1];
    }

    cfg->energyabs = cfg->energytot - cfg->energyesc;

    /* trajectory records, appended to the caller's (src/mcx_host.cpp:1173-1193);
       the counter runs over all repetitions, so each record is read once */
    if (g.savetraj) {
        uint32_t njump = 0;
        HIP_TRY(hipMemcpy(&njump, djump, sizeof(uint32_t), hipMemcpyDeviceToHost));

        if (njump > cfg->maxjumpdebug) {
            fprintf(cfg->flog ? cfg->flog : stdout,
                    "WARNING: the saved trajectory positions are more than what your have specified (%u > %u), please "
                    "use the --maxjumpdebug option to specify a greater number\n", njump, cfg->maxjumpdebug);
            njump = cfg->maxjumpdebug;
        }

        if (njump) {
            float* buf = (float*)realloc(cfg->exportdebugdata, ((size_t)cfg->debugdatalen + njump) *
                                         g.trajlen * sizeof(float));

            if (!buf) {
                HIP_TRY(hipErrorOutOfMemory);
            }

            cfg->exportdebugdata = buf;
            HIP_TRY(hipMemcpy(buf + (size_t)cfg->debugdatalen * g.trajlen, ddebug,
                              (size_t)njump * g.trajlen * sizeof(float), hipMemcpyDeviceToHost));
            cfg->debugdatalen += njump;
        }
    }

    /* detected photons (src/mcx_host.cpp:1195-1237) */
    if (cfg->issavedet) {
        HIP_TRY(hipMemcpy(&hdetcount, ddetcount, sizeof(uint32_t), hipMemcpyDeviceToHost));

        if (cfg->issaveref > 1) {
            hdetcount = cfg->maxdetphoton;   /* every voxel's exit record (src/mcx_core.cl:3304-3306) */
        }

        cfg->his.detected = hdetcount;

        if (hdetcount > cfg->maxdetphoton) {
            fprintf(cfg->flog ? cfg->flog : stdout,
                    "WARNING: the detected photon number is more than what your have specified (%u > %u), please use "
                    "the -H option to specify a greater number\n", hdetcount, cfg->maxdetphoton);
        }

        uint32_t nsave = hdetcount < cfg->maxdetphoton ? hdetcount : cfg->maxdetphoton;
        cfg->exportdetected = (float*)realloc(cfg->exportdetected, (size_t)(nsave ? nsave : 1) * reclen * sizeof(float));

        if (nsave && reclen) {
            HIP_TRY(hipMemcpy(cfg->exportdetected, ddetrec, (size_t)nsave * reclen * sizeof(float), hipMemcpyDeviceToHost));
        }

        if (cfg->issaveseed) {
            cfg->seeddata = realloc(cfg->seeddata, (size_t)(nsave ? nsave : 1) * 16);

            if (nsave) {
                HIP_TRY(hipMemcpy(cfg->seeddata, ddetseed, (size_t)nsave * 16, hipMemcpyDeviceToHost));
            }
        }

        if (ddetw && nsave) {
            HIP_TRY(hipMemcpy(cfg->exportdetweight, ddetw, (size_t)nsave * sizeof(float), hipMemcpyDeviceToHost));
        }

        cfg->detectedcount = nsave;
    }

    /* fluence: double accumulator -> float, normalised on the device
       (src/mcx_host.cpp:1382-1465, src/mcx_utils.c:1208-1218) */
    if (cfg->issave2pt) {
        float scale = 1.f;
        HIP_TRY(hipEventRecord(ev1, stream));
        int normalize = cfg->isnormalized && !g.debugrng;

        if (normalize) {
            scale = mcx_output_scale(cfg);

            if (scale != 0.f) {
                cfg->normalizer = scale;
                cfg->his.normalizer = scale;
            }
        }

        if (wantfloat) {
            /* shared patterns are scaled per slot on the host after the read-back */
            HIP_TRY(mcx_launch_finalize(dfield, dout, fieldlen, g.nrep,
                                        (normalize && scale != 0.f && g.npattern <= 1) ? scale : 1.f, stream));
        } else if (g.nrep > 1) {
            /* fold the replicas into replica 0 without a float output */
            HIP_TRY(mcx_launch_finalize(dfield, NULL, fieldlen, g.nrep, 1.f, stream));
        }

        /* adjoint outputs (src/mcx_host.cpp:1467-1642): products of the normalised
           per-source fluences of each user source s and detector-as-source d,
           formed on the device from the float field, scaled by -Vvox (mua) or
           -unitinmm (D, with 1/(3(1-g)mus^2) or 1/(3(1-g)^2 mus^2) per label for
           mus and mus'), into cfg->exportjacob [Nvox, Ns*Nd] (a second block for
           the two-output types) */
        if (wantadj) {
            const uint32_t ns = g.adjsrc, nd = cfg->detnum;
            const int dual = adjdual, musp = adjmusp, needop = adjop;
            const float Vvox = cfg->steps.x * cfg->steps.y * cfg->steps.z;
            float hop[256];

            for (int m = 0; m < 256; m++) {
                hop[m] = 0.f;

                if (needop && (unsigned int)m < cfg->medianum) {
                    const float mus = cfg->prop[m].mus, onemg = 1.f - cfg->prop[m].g;

                    if (mus > 0.f && onemg > 0.f) {
                        hop[m] = musp ? 1.f / (3.f * onemg * onemg * mus * mus) : 1.f / (3.f * onemg * mus * mus);
                    }
                }
            }

            if (needop) {
                HIP_TRY(hipMemcpyAsync(dopscale, hop, sizeof(hop), hipMemcpyHostToDevice, stream));
            }

            /* RF forward: complex products of the real and imaginary halves; the
               layout is [re] [im] for one output, [re1] [re2] [im1] [im2] for two
               (src/mcx_host.cpp:1559-1568) */
            const float* fim = g.rfforward ? dout + g.imoff : NULL;
            const size_t imblock = g.rfforward ? adjlen * (dual ? 2 : 1) : 0;

            if (dual || cfg->outputtype == otAdjoint) {
                HIP_TRY(mcx_launch_adjoint(dout, fim, djacob, fim ? djacob + imblock : NULL, (uint32_t)nvox, cfg->dim.x,
                                           cfg->dim.y, maxgate, ns, nd, 0, -Vvox, NULL, dmedia, stream));
            }

            if (dual || cfg->outputtype != otAdjoint) {
                float* o = djacob + (dual ? adjlen : 0);
                HIP_TRY(mcx_launch_adjoint(dout, fim, o, fim ? o + imblock : NULL, (uint32_t)nvox, cfg->dim.x,
                                           cfg->dim.y, maxgate, ns, nd, 1, -cfg->unitinmm, needop ? dopscale : NULL,
                                           dmedia, stream));
            }
        }

        if (cfg->exportfield32) {
            /* the replicas are folded by now: a plain float64 -> float32 pass */
            HIP_TRY(mcx_launch_finalize(dfield, draw32, fieldlen, 1, 1.f, stream));
        }

        HIP_TRY(hipEventRecord(ev2, stream));

        if (wantadj) {
            free(cfg->exportjacob);
            const size_t jlen = adjlen * (adjdual ? 2 : 1) * (g.rfforward ? 2 : 1);
            cfg->exportjacob = (float*)malloc(jlen * sizeof(float));

            if (!cfg->exportjacob) {
                HIP_TRY(hipErrorOutOfMemory);
            }

            HIP_TRY(hipMemcpyAsync(cfg->exportjacob, djacob, jlen * sizeof(float), hipMemcpyDeviceToHost, stream));
        }

        if (cfg->exportfield32) {
            HIP_TRY(hipMemcpyAsync(cfg->exportfield32, draw32, fieldlen * sizeof(float), hipMemcpyDeviceToHost,
                                   stream));
        }

        if (cfg->exportfield64) {
            HIP_TRY(hipMemcpyAsync(cfg->exportfield64, dfield, fieldlen * sizeof(double), hipMemcpyDeviceToHost,
                                   stream));
        }

        if (wantfloat) {
            hfield = (float*)malloc(fieldlen * sizeof(float));

            if (!hfield) {
                HIP_TRY(hipErrorOutOfMemory);
            }

            HIP_TRY(hipMemcpyAsync(hfield, dout, fieldlen * sizeof(float), hipMemcpyDeviceToHost, stream));
        }

        HIP_TRY(hipStreamSynchronize(stream));
        {
            float postms = 0.f;
            HIP_TRY(hipEventElapsedTime(&postms, ev1, ev2));
            cfg->jobms = kernelms + postms;
        }

        if (wantfloat && normalize && scale == 0.f && cfg->seed == SEED_FROM_FILE && cfg->replaydet == -1) {
            /* one slice per detector (src/mcx_host.cpp:1398-1422), and with srcid < 0 one
               per (source, detector) pair, each scaled by its own photons' weights */
            const int nsrc = (cfg->extrasrclen && cfg->srcid < 0) ? (int)cfg->extrasrclen + 1 : 1;

            for (int s = 1; s <= nsrc; s++) {
                for (int d = 1; d <= (int)cfg->detnum; d++) {
                    float sd = 0.f;

                    for (size_t i = 0; i < cfg->nphoton; i++)
                        if ((cfg->replay.detid[i] & 0xFFFF) == d && (nsrc == 1 || (cfg->replay.detid[i] >> 16) == s)) {
                            sd += cfg->replay.weight[i];
                        }

                    if (sd > 0.f) {
                        sd = cfg->unitinmm / sd;
                    }

                    for (size_t part = 0; part < (g.imoff ? 2u : 1u); part++) {
                        mcx_normalize(hfield + part * g.imoff + ((size_t)(s - 1) * cfg->detnum + (d - 1)) * nvox * maxgate,
                                      sd, (int)(nvox * maxgate), 1, 0, 1);
                    }
                }
            }
        }

        if (wantfloat && normalize && scale != 0.f && g.npattern > 1) {
            /* pattern k: npixel / (its summed weights) x the unit-photon scale
               (src/mcx_host.cpp:1351-1364, :1436-1447), Kahan sums in float */
            const int psize = (int)mcx_pattern_pixels(cfg);

            for (uint32_t k = 0; k < g.npattern; k++) {
                float sum = 0.f, c = 0.f;

                for (int j = 0; j < psize; j++) {
                    mcx_kahanSum(&sum, &c, cfg->srcpattern[(size_t)j * g.npattern + k]);
                }

                mcx_normalize(hfield, (float)psize / sum * scale, (int)(fieldlen / g.npattern), cfg->isnormalized, (int)k,
                              (int)g.npattern);
            }
        }

        if (wantfloat) {
            if (!cfg->exportfield) {
                cfg->exportfield = (float*)calloc(fieldlen * 2, sizeof(float));
            }

            for (size_t i = 0; i < fieldlen; i++) {
                cfg->exportfield[i] += hfield[i];
            }
        }
    }

    /* a stand-alone caller gets its output files here (src/mcx_host.cpp:1646-1664);
       the MATLAB/Python bindings take the buffers instead */
    if (cfg->parentid == mpStandalone) {
        if (cfg->issave2pt && cfg->exportfield) {
            mcx_savedata(cfg->exportfield, fieldlen, cfg);
        }

        if (cfg->issavedet && cfg->exportdetected) {
            cfg->his.unitinmm = cfg->unitinmm;
            cfg->his.savedphoton = cfg->detectedcount;
            cfg->his.totalphoton = (unsigned int)cfg->nphoton;
            cfg->his.seedbyte = (cfg->issaveseed && cfg->seeddata) ? 16 : 0;
            cfg->his.detected = cfg->detectedcount;
            mcx_savedetphoton(cfg->exportdetected, cfg->seeddata, (int)cfg->detectedcount, 0, cfg);
        }

        /* trajectories: a .mct file in the .mch layout (src/mcx_host.cpp:1666-1672), 6 columns or
           10 with Stokes vectors (the reference's header says 6 for those too) */
        if (g.savetraj && cfg->exportdebugdata) {
            const unsigned int colcount = cfg->his.colcount, seedbyte = cfg->his.seedbyte;
            cfg->his.colcount = g.trajlen;
            cfg->his.savedphoton = cfg->debugdatalen;
            cfg->his.totalphoton = (unsigned int)cfg->nphoton;
            cfg->his.detected = 0;
            cfg->his.seedbyte = 0;
            mcx_savedetphoton(cfg->exportdebugdata, NULL, (int)cfg->debugdatalen, 0, cfg);
            cfg->his.colcount = colcount;
            cfg->his.seedbyte = seedbyte;
        }
    }

    free(hseed);
    free(hgenergy);
    free(hfield);
    free(hpolcdf);
    bufs.release();
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    (void)hipEventDestroy(ev2);
    (void)hipStreamDestroy(stream);
    return;

fail:
    free(hseed);
    free(hgenergy);
    free(hfield);
    free(hpolcdf);
    bufs.release();

    if (ev0) {
        (void)hipEventDestroy(ev0);
    }

    if (ev2) {
        (void)hipEventDestroy(ev2);
    }

    if (ev1) {
        (void)hipEventDestroy(ev1);
    }

    if (stream) {
        (void)hipStreamDestroy(stream);
    }

    snprintf(msg, sizeof(msg), "HIP error %d (%s) at %s:%d", (int)err, hipGetErrorString(err), __FILE__, errline);
    mcx_error(-(int)err, msg, __FILE__, errline);
}

extern "C" int mcx_try_run_simulation(Config* cfg) {
    jmp_buf jb;
    void* old = mcx_push_errjmp(&jb);
    int rc = setjmp(jb);

    if (rc == 0) {
        mcx_run_simulation(cfg, NULL, NULL);
    }

    mcx_pop_errjmp(old);
    return rc;
}

/* ---------------- device self-checks ---------------- */
extern "C" int mcx_device_fpmath(int gpuid, int which, int mode, const float* x, float* y, size_t n) {
    float *dx = NULL, *dy = NULL;
    size_t nin = (which == 6) ? 2 * n : n;
    hipError_t e = hipSetDevice(gpuid);

    if (e == hipSuccess) {
        e = hipMalloc((void**)&dx, nin * sizeof(float));
    }

    if (e == hipSuccess) {
        e = hipMalloc((void**)&dy, n * sizeof(float));
    }

    if (e == hipSuccess) {
        e = hipMemcpy(dx, x, nin * sizeof(float), hipMemcpyHostToDevice);
    }

    if (e == hipSuccess) {
        e = mode ? mcx_launch_fpmath_parity(which, dx, dy, n, 0) : mcx_launch_fpmath_fast(which, dx, dy, n, 0);
    }

    if (e == hipSuccess) {
        e = hipMemcpy(y, dy, n * sizeof(float), hipMemcpyDeviceToHost);
    }

    if (dx) {
        (void)hipFree(dx);
    }

    if (dy) {
        (void)hipFree(dy);
    }

    return (int)e;
}

extern "C" int mcx_device_rng_stream(int gpuid, const unsigned int seed4[4], unsigned int n, float* out) {
    uint32_t* ds = NULL;
    float* dout = NULL;
    hipError_t e = hipSetDevice(gpuid);

    if (e == hipSuccess) {
        e = hipMalloc((void**)&ds, 4 * sizeof(uint32_t));
    }

    if (e == hipSuccess) {
        e = hipMalloc((void**)&dout, n * sizeof(float));
    }

    if (e == hipSuccess) {
        e = hipMemcpy(ds, seed4, 4 * sizeof(uint32_t), hipMemcpyHostToDevice);
    }

    if (e == hipSuccess) {
        e = mcx_launch_rngstream(ds, n, dout, 0);
    }

    if (e == hipSuccess) {
        e = hipMemcpy(out, dout, n * sizeof(float), hipMemcpyDeviceToHost);
    }

    if (ds) {
        (void)hipFree(ds);
    }

    if (dout) {
        (void)hipFree(dout);
    }

    return (int)e;
}

extern "C" int mcx_device_atomic_ceiling(int gpuid, unsigned int ncell, double* rate) {
    double* grid = NULL;
    hipEvent_t e0 = NULL, e1 = NULL;
    const uint32_t nblock = 2560, k = 64;
    float ms = 0.f, best = 1e30f;
    hipError_t e = hipSetDevice(gpuid);

    *rate = 0.0;

    if (ncell == 0) {
        return (int)hipErrorInvalidValue;
    }

    if (e == hipSuccess) {
        e = hipMalloc((void**)&grid, (size_t)ncell * sizeof(double));
    }

    if (e == hipSuccess) {
        e = hipEventCreate(&e0);
    }

    if (e == hipSuccess) {
        e = hipEventCreate(&e1);
    }

    for (int rep = 0; rep < 3 && e == hipSuccess; rep++) {
        e = hipMemset(grid, 0, (size_t)ncell * sizeof(double));

        if (e == hipSuccess) {
            e = hipEventRecord(e0, 0);
        }

        if (e == hipSuccess) {
            e = mcx_launch_atomic_probe(grid, ncell, nblock, k, 0);
        }

        if (e == hipSuccess) {
            e = hipEventRecord(e1, 0);
        }

        if (e == hipSuccess) {
            e = hipEventSynchronize(e1);
        }

        if (e == hipSuccess) {
            e = hipEventElapsedTime(&ms, e0, e1);
        }

        if (e == hipSuccess && ms < best) {
            best = ms;
        }
    }

    if (e == hipSuccess && best > 0.f) {
        *rate = (double)nblock * 256.0 * k / (best * 1e-3);
    }

    if (e0) {
        (void)hipEventDestroy(e0);
    }

    if (e1) {
        (void)hipEventDestroy(e1);
    }

    if (grid) {
        (void)hipFree(grid);
    }

    return (int)e;
}

/* ---------------- no-allocation entry point ----------------
   mcx_run_simulation_noalloc keeps its per-run device state in `scratch`, a
   file-scope array that each device holds for the life of the process: the
   per-lane seed words and energy tallies and the staged medium, phase-function
   and source tables.  Host code never uses the array's name as an address: on
   a device's first call it asks hipGetSymbolAddress for that device's copy and
   keeps the pointer (mcx_noalloc_base).  The caller passes the volume and the
   outputs in device memory; nothing is allocated per call, on the device or
   on the host (seed words and energy tallies pass through fixed host chunks). */
#define MCX_NOALLOC_LANES ((size_t)262144) /* one resident wave: 256 CUs x 1024 lanes */
#define MCX_NOALLOC_MEDIA ((size_t)4096)
#define MCX_NOALLOC_PHASE ((size_t)4096)
#define MCX_NOALLOC_SRC ((size_t)256)
#define MCX_NOALLOC_OFF_ENERGY (MCX_NOALLOC_LANES * 4 * sizeof(uint32_t))
#define MCX_NOALLOC_OFF_PROP (MCX_NOALLOC_OFF_ENERGY + MCX_NOALLOC_LANES * 2 * sizeof(float))
#define MCX_NOALLOC_OFF_PHASE (MCX_NOALLOC_OFF_PROP + MCX_NOALLOC_MEDIA * sizeof(float4))
#define MCX_NOALLOC_OFF_SRC (MCX_NOALLOC_OFF_PHASE + MCX_NOALLOC_PHASE * sizeof(float))
#define MCX_NOALLOC_BYTES (MCX_NOALLOC_OFF_SRC + MCX_NOALLOC_SRC * sizeof(ExtraSrc))
#define MCX_NOALLOC_CHUNK ((size_t)65536)

static __device__ __align__(256) unsigned char scratch[MCX_NOALLOC_BYTES];

static void* mcx_noalloc_base[MAX_DEVICE];
static uint32_t mcx_noalloc_words[MCX_NOALLOC_CHUNK];
static float mcx_noalloc_energy[MCX_NOALLOC_CHUNK];

/* zero the energy tallies and the caller's float64 field (a fill kernel rather
   than hipMemsetAsync, so the call can sit in a captured stream) */
static __global__ void noalloc_clear_kernel(float* __restrict__ energy, size_t nenergy, double* __restrict__ field,
                                     size_t nfield) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;

    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nenergy; i += stride) {
        energy[i] = 0.f;
    }

    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nfield; i += stride) {
        field[i] = 0.0;
    }
}

static void noalloc_run(Config* cfg, const uint32_t* media, size_t nmedia, double* field64, float* fluence,
                        size_t fieldlen, hipStream_t stream) {
    hipError_t err = hipSuccess;
    int errline = 0;
    char msg[512];
    int dev = 0, ncu = 0;
    GlibcRand seedgen(cfg->seed > 0 ? (unsigned int)cfg->seed : (unsigned int)time(0));

    mcx_check_supported(cfg);

    if (cfg->issavedet || cfg->seed == SEED_FROM_FILE || MCX_IS_ADJOINT_TYPE(cfg->outputtype) ||
            (cfg->debuglevel & (MCX_DEBUG_RNG | MCX_DEBUG_MOVE | MCX_DEBUG_MOVE_ONLY)) || cfg->muamus ||
            cfg->mediabyte > 4 || cfg->omega > 0.f || !cfg->issave2pt || cfg->polmedianum) {
        MCX_ERROR(-4, "mcx_run_simulation_noalloc runs forward fluence simulations only: no detected photons, "
                  "replay, adjoint outputs, RNG dump, trajectories, RF or per-voxel media, and the field must be saved");
    }

    if (!media || !field64) {
        MCX_ERROR(-4, "mcx_run_simulation_noalloc needs the volume and the float64 field in device memory");
    }

    if (cfg->srctype == MCX_SRC_PATTERN || cfg->srctype == MCX_SRC_PATTERN3D || cfg->nangle) {
        MCX_ERROR(-4, "mcx_run_simulation_noalloc does not stage pattern or launch-angle tables: use "
                  "mcx_run_simulation for pattern sources and angle-CDF launches");
    }

    if (cfg->medianum > MCX_NOALLOC_MEDIA || cfg->nphase > MCX_NOALLOC_PHASE || cfg->extrasrclen > MCX_NOALLOC_SRC) {
        MCX_ERROR(-4, "mcx_run_simulation_noalloc holds at most 4096 media, 4096 phase-function samples and 256 "
                  "extra sources");
    }

    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICE ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0) {
        MCX_ERROR(-99, "mcx_run_simulation_noalloc: no usable current HIP device");
    }

    if (!mcx_noalloc_base[dev]) {
        void* p = NULL;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));

        if (e != hipSuccess || !p) {
            snprintf(msg, sizeof(msg), "mcx_run_simulation_noalloc: scratch address of device %d not found (%s)", dev,
                     hipGetErrorString(e));
            MCX_ERROR(-99, msg);
        }

        mcx_noalloc_base[dev] = p;
    }

    unsigned char* base = (unsigned char*)mcx_noalloc_base[dev];
    size_t nblocksize = cfg->autopilot ? 64 : cfg->nblocksize;
    size_t nthread = cfg->autopilot ? (size_t)2560 * (size_t)ncu : cfg->nthread;

    if (nblocksize == 0 || nthread < nblocksize) {
        MCX_ERROR(-4, "thread number must be at least one block");
    }

    nthread = (nthread / nblocksize) * nblocksize;
    size_t lblock = 0, nlaunch = 0, shmem = 0;
    int reflect = 0;
    MCXKParam g;
    mcx_plan_run(cfg, dev, ncu, &nthread, &g, &lblock, &nlaunch, &shmem, &reflect);

    if (nthread > MCX_NOALLOC_LANES) {
        snprintf(msg, sizeof(msg), "mcx_run_simulation_noalloc: %zu lanes exceed its scratch (%zu lanes); set "
                 "nthread (autopilot 0) or use fast arithmetic", nthread, MCX_NOALLOC_LANES);
        MCX_ERROR(-4, msg);
    }

    if (nmedia != g.nvox || fieldlen != g.fieldlen) {
        snprintf(msg, sizeof(msg), "mcx_run_simulation_noalloc: buffer sizes %zu/%zu do not match the volume (%u "
                 "voxels) and the field (%u values)", nmedia, fieldlen, g.nvox, g.fieldlen);
        MCX_ERROR(-4, msg);
    }

    uint32_t* dseed = (uint32_t*)base;
    float* denergy = (float*)(base + MCX_NOALLOC_OFF_ENERGY);
    float4* dprop = (float4*)(base + MCX_NOALLOC_OFF_PROP);
    float* dphase = (float*)(base + MCX_NOALLOC_OFF_PHASE);
    float4* dsrc = (float4*)(base + MCX_NOALLOC_OFF_SRC);
    const unsigned int respin = (unsigned int)(cfg->respin < 0 ? -cfg->respin : (cfg->respin ? cfg->respin : 1));
    double t0 = wall_ms();

    g.nrep = 1;
    g.media = media;
    g.prop = dprop;
    g.detpos = NULL;
    g.invcdf = cfg->nphase ? dphase : NULL;
    g.srcdata = cfg->extrasrclen ? dsrc : NULL;
    g.nseed = dseed;
    g.field = field64;
    g.genergy = denergy;
    g.detcount = NULL;

    cfg->detectedcount = 0;
    cfg->his.detected = 0;
    cfg->his.respin = cfg->respin;
    cfg->his.colcount = g.reclen;
    cfg->energytot = 0;
    cfg->energyesc = 0;

    HIP_TRY(hipMemcpyAsync(dprop, cfg->prop, cfg->medianum * sizeof(float4), hipMemcpyHostToDevice, stream));

    if (cfg->nphase) {
        HIP_TRY(hipMemcpyAsync(dphase, cfg->invcdf, cfg->nphase * sizeof(float), hipMemcpyHostToDevice, stream));
    }

    if (cfg->extrasrclen) {
        HIP_TRY(hipMemcpyAsync(dsrc, cfg->srcdata, cfg->extrasrclen * sizeof(ExtraSrc), hipMemcpyHostToDevice, stream));
    }

    {
        int nb = (int)((fieldlen > 2 * nthread ? fieldlen : 2 * nthread) / 256 + 1);
        nb = nb > 4096 ? 4096 : nb;
        hipLaunchKernelGGL(noalloc_clear_kernel, dim3(nb), dim3(256), 0, stream, denergy, 2 * nthread, field64,
                           fieldlen);
        HIP_TRY(hipGetLastError());
    }

    for (unsigned int pass = 0; pass < respin; pass++) {
        /* seeds: the next nthread*4 words of the srand(seed) stream, as
           mcx_run_simulation draws them, through the fixed host chunk */
        for (size_t off = 0; off < nthread * 4; off += MCX_NOALLOC_CHUNK) {
            const size_t n = nthread * 4 - off < MCX_NOALLOC_CHUNK ? nthread * 4 - off : MCX_NOALLOC_CHUNK;
            seedgen.fill(mcx_noalloc_words, n);
            HIP_TRY(hipMemcpyAsync(dseed + off, mcx_noalloc_words, n * sizeof(uint32_t), hipMemcpyHostToDevice,
                                   stream));
            HIP_TRY(hipStreamSynchronize(stream));
        }

        if (cfg->mathmode == 1) {
            HIP_TRY(mcx_launch_parity(&g, (int)nlaunch, (int)lblock, shmem, stream, 0, reflect));
        } else {
            HIP_TRY(mcx_launch_fast(&g, (int)nlaunch, (int)lblock, shmem, stream, 0, reflect));
        }
    }

    /* energy tallies summed in thread order in double, as mcx_run_simulation does */
    for (size_t off = 0; off < nthread * 2; off += MCX_NOALLOC_CHUNK) {
        const size_t n = nthread * 2 - off < MCX_NOALLOC_CHUNK ? nthread * 2 - off : MCX_NOALLOC_CHUNK;
        HIP_TRY(hipMemcpyAsync(mcx_noalloc_energy, denergy + off, n * sizeof(float), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));

        for (size_t i = 0; i < n; i += 2) {
            cfg->energyesc += mcx_noalloc_energy[i];
            cfg->energytot += mcx_noalloc_energy[i + 1];
        }
    }

    cfg->energyabs = cfg->energytot - cfg->energyesc;

    if (fluence) {
        float scale = 1.f;

        if (cfg->isnormalized) {
            scale = mcx_output_scale(cfg);
            cfg->normalizer = scale;
            cfg->his.normalizer = scale;
        }

        HIP_TRY(mcx_launch_finalize(field64, fluence, fieldlen, 1, scale, stream));
        HIP_TRY(hipStreamSynchronize(stream));
    }

    cfg->runtimems = wall_ms() - t0;
    cfg->runtime = (unsigned int)(cfg->runtimems + 0.5);
    cfg->kernelms = 0.f;
    cfg->jobms = 0.f;
    return;

fail:
    snprintf(msg, sizeof(msg), "HIP error %d (%s) at %s:%d", (int)err, hipGetErrorString(err), __FILE__, errline);
    mcx_error(-(int)err, msg, __FILE__, errline);
}

extern "C" int mcx_run_simulation_noalloc(Config* cfg, const unsigned int* media, size_t nmedia, double* field64,
        float* fluence, size_t fieldlen, void* stream) {
    jmp_buf jb;
    void* old = mcx_push_errjmp(&jb);
    int rc = setjmp(jb);

    if (rc == 0) {
        noalloc_run(cfg, media, nmedia, field64, fluence, fieldlen, (hipStream_t)stream);
    }

    mcx_pop_errjmp(old);
    return rc;
}

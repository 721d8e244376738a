/*
 * mcx_batch.hip - many independent simulations in one persistent launch.
 *
 * The reference's host runs one Config per mcx_run_simulation call; a study of
 * many small domains then pays one launch, one set of uploads and one tail of
 * idle lanes per simulation.  Here n prepared Configs (own volume, media,
 * source, photon count, time gates) are staged once and each step runs them
 * all in one persistent kernel (batch_kernel in mcx_core_impl.h):
 *   - lanes draw each instance's photons one at a time from a counter in
 *     global memory, so workgroups share an instance without item barriers;
 *   - instances are split into one list per XCD, balanced by predicted cost
 *     (longest processing time first) and run heaviest first, so an XCD's L2
 *     holds few volumes and grids at a time;
 *   - a step is prepare (zero grids and tallies, restore the counters) +
 *     transport + finalize (per-instance normalisation on the device),
 *     optionally captured once as a HIP graph and replayed.
 * The schedule is tuned once: the launch shape that keeps the most lanes
 * resident, and per-instance costs measured by a probe step at a scaled-down
 * photon count (the kernel's wall-clock per instance) for the list balance.
 *
 * Device memory of a batch: per instance its labels (uint32 nvox), media
 * table, phase table, extra sources, float64 grid and float32 output; shared:
 * kernel parameters, normalisation records, the instance lists and counters,
 * one RNG state per lane (continued from step to step) and two energy tallies
 * per instance.
 */
#include <math.h>
#include <setjmp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <hip/hip_runtime.h>

#include "mcx_core.h"
#include "mcx_host.h"

extern "C" void* mcx_push_errjmp(void* jb);
extern "C" void mcx_pop_errjmp(void* old);
void mcx_check_supported(const Config* cfg);
size_t mcx_kparam_setup(Config* cfg, MCXKParam* gp, size_t nthread, size_t lblock, int* reflect, size_t boxbytes);

#define MCX_BATCH_MAXBUF 16
#define MCX_BATCH_MAXBLOCK ((size_t)MCX_BATCH_BLOCK_MAX)
/* LDS kept back from the deposit-box budget (static LDS, allocation granularity)
   and the per-workgroup LDS cap (gfx950: 160 KB) */
#define MCX_BATCH_LDS_SLACK 1024u
#define MCX_BATCH_LDS_LIMIT (160u * 1024u)
#define MCX_BATCH_TUNE_PHOTONS 1.0e6
#define MCX_BATCH_TUNE_MS 150.0

struct MCXBatch {
    int dev;
    unsigned int n;
    Config** cfg;
    MCXBatchOpt* opt;
    hipStream_t stream;
    hipEvent_t ev0, ev1, ev2, ev3;
    void* buf[MCX_BATCH_MAXBUF];
    int nbuf;
    /* per instance */
    MCXKParam* prm;                 /* host copy of the kernel parameters */
    MCXBatchNorm* norm;
    size_t* fieldoff;               /* offset of each grid in dfield / output in dout */
    uint32_t** dmedia;
    float4** dprop;
    float** dinvcdf;
    float4** dsrcdata;
    /* shared device buffers */
    MCXKParam* dprm;
    MCXBatchNorm* dnorm;
    uint32_t* dnseed;
    double* denergy;
    double* dfield;
    float* dout;
    size_t nfield, maxlen, maxlanes;
    /* schedule */
    uint32_t* qlist;                /* instances of list 0, then list 1, ... */
    uint32_t* state0;               /* initial {head, end} per list, {0, photons} per instance */
    uint32_t* qof;                  /* list of each instance */
    uint32_t* order;                /* instances by decreasing predicted cost */
    double* weight;                 /* predicted cost of each instance */
    uint32_t* dqlist;
    uint32_t* dstate;
    uint32_t* dstate0;
    unsigned long long* dcost;      /* per-instance workgroup time of a probe step */
    uint32_t nqueue, nstate;
    size_t lblock, nblock, shmem;
    int reflect;
    hipGraph_t graph;
    hipGraphExec_t exec;
};

#define BHIP(call)                                                                    \
    do {                                                                              \
        hipError_t e__ = (call);                                                      \
        if (e__ != hipSuccess) {                                                      \
            char m__[256];                                                            \
            snprintf(m__, sizeof(m__), "HIP error %d (%s) in the batched launch",     \
                     (int)e__, hipGetErrorString(e__));                               \
            mcx_error(-(int)e__, m__, __FILE__, __LINE__);                            \
        }                                                                             \
    } while (0)

template <typename T>
static T* batch_alloc(MCXBatch* b, size_t n) {
    void* p = NULL;

    if (b->nbuf >= MCX_BATCH_MAXBUF) {
        MCX_ERROR(-4, "too many device buffers in a batch");
    }

    BHIP(hipMalloc(&p, (n ? n : 1) * sizeof(T)));
    b->buf[b->nbuf++] = p;
    return (T*)p;
}

static void* host_calloc(size_t n, size_t size) {
    void* p = calloc(n ? n : 1, size);

    if (!p) {
        MCX_ERROR(-4, "out of host memory");
    }

    return p;
}

static void batch_free(MCXBatch* b) {
    if (!b) {
        return;
    }

    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
    }

    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
    }

    for (int i = 0; i < b->nbuf; i++) {
        (void)hipFree(b->buf[i]);
    }

    /* per-instance buffers share one allocation each kind; freed above */
    hipEvent_t ev[4] = {b->ev0, b->ev1, b->ev2, b->ev3};

    for (int i = 0; i < 4; i++) {
        if (ev[i]) {
            (void)hipEventDestroy(ev[i]);
        }
    }

    if (b->stream) {
        (void)hipStreamDestroy(b->stream);
    }

    free(b->prm);
    free(b->norm);
    free(b->fieldoff);
    free(b->dmedia);
    free(b->dprop);
    free(b->dinvcdf);
    free(b->dsrcdata);
    free(b->qlist);
    free(b->state0);
    free(b->qof);
    free(b->order);
    free(b->weight);
    free(b->cfg);
    free(b);
}

/* kernel parameters of every instance for workgroups of lblock lanes, each
   with an LDS deposit box of at most boxbytes; returns the LDS bytes the launch
   needs (the largest instance's) */
static size_t batch_params(MCXBatch* b, size_t lblock, size_t boxbytes) {
    size_t shmem = 0;
    b->reflect = 0;

    for (unsigned int i = 0; i < b->n; i++) {
        Config* cfg = b->cfg[i];
        MCXKParam& g = b->prm[i];
        int reflect = 0;
        size_t s = mcx_kparam_setup(cfg, &g, lblock, lblock, &reflect, boxbytes);
        shmem = s > shmem ? s : shmem;
        b->reflect |= reflect;
        g.nrep = 1;
        g.dynamic = 1;
        g.media = b->dmedia[i];
        g.prop = b->dprop[i];
        g.invcdf = b->dinvcdf[i];
        g.srcdata = b->dsrcdata[i];
        g.field = b->dfield + b->fieldoff[i];
        g.genergy = NULL;
        g.nseed = NULL;
    }

    return shmem;
}

/* the initial schedule state: per list {head, end} into qlist, per instance
   {0, photons}, photon counts scaled by frac (at least one photon) */
static size_t batch_state(MCXBatch* b, double frac) {
    uint32_t k = 0;
    size_t total = 0;
    uint32_t* pool = b->state0 + 2 * b->nqueue;

    for (uint32_t q = 0; q < b->nqueue; q++) {
        b->state0[2 * q] = k;

        for (unsigned int r = 0; r < b->n; r++) {
            const unsigned int i = b->order[r];

            if (b->qof[i] == q) {
                b->qlist[k++] = i;
            }
        }

        b->state0[2 * q + 1] = k;
    }

    for (unsigned int i = 0; i < b->n; i++) {
        size_t np = (size_t)(b->cfg[i]->nphoton * frac + 0.5);
        np = (np == 0 && b->cfg[i]->nphoton) ? 1 : np;
        pool[2 * i] = 0;
        pool[2 * i + 1] = (uint32_t)np;
        total += np;
    }

    return total;
}

/* instances to lists by predicted cost (b->weight), heaviest first onto the
   least-loaded list; a list runs its instances heaviest first */
static void batch_assign(MCXBatch* b) {
    uint32_t* order = b->order;
    double* load = (double*)host_calloc(b->nqueue, sizeof(double));

    for (unsigned int i = 0; i < b->n; i++) {
        order[i] = i;
    }

    for (unsigned int i = 1; i < b->n; i++) {   /* insertion sort, decreasing weight */
        uint32_t v = order[i];
        unsigned int j = i;

        while (j > 0 && b->weight[order[j - 1]] < b->weight[v]) {
            order[j] = order[j - 1];
            j--;
        }

        order[j] = v;
    }

    for (unsigned int r = 0; r < b->n; r++) {
        unsigned int i = order[r];
        uint32_t best = 0;

        for (uint32_t q = 1; q < b->nqueue; q++) {
            best = load[q] < load[best] ? q : best;
        }

        b->qof[i] = best;
        load[best] += b->weight[i];
    }

    free(load);
}

static void batch_upload_schedule(MCXBatch* b) {
    BHIP(hipMemcpyAsync(b->dprm, b->prm, b->n * sizeof(MCXKParam), hipMemcpyHostToDevice, b->stream));
    BHIP(hipMemcpyAsync(b->dqlist, b->qlist, b->n * sizeof(uint32_t), hipMemcpyHostToDevice, b->stream));
    BHIP(hipMemcpyAsync(b->dstate0, b->state0, b->nstate * sizeof(uint32_t), hipMemcpyHostToDevice, b->stream));
    BHIP(hipStreamSynchronize(b->stream));
}

/* one step on the stream: prepare, transport, finalize; ev1/ev2 bracket the
   transport kernel when timed; probe also records each instance's cost */
static void batch_enqueue(MCXBatch* b, int timed, int probe) {
    BHIP(mcx_launch_batch_prep(b->dfield, b->nfield, b->denergy, 2 * b->n, b->dstate, b->dstate0, b->nstate,
                               b->stream));

    if (probe) {
        BHIP(hipMemsetAsync(b->dcost, 0, b->n * sizeof(unsigned long long), b->stream));
    }

    if (timed) {
        BHIP(hipEventRecord(b->ev1, b->stream));
    }

    BHIP(mcx_launch_batch(b->dprm, b->dqlist, b->nqueue, b->dstate, b->dnseed, b->denergy, probe ? b->dcost : NULL,
                          (int)b->nblock, (int)b->lblock, b->shmem, b->reflect, b->stream));

    if (timed) {
        BHIP(hipEventRecord(b->ev2, b->stream));
    }

    BHIP(mcx_launch_batch_finalize(b->dnorm, b->denergy, b->n, b->maxlen, b->stream));
}

/* resident workgroups for lblock lanes; sets b->lblock, b->nblock, b->shmem.
   The deposit box of each instance takes the LDS a CU has left for each of the
   workgroups its registers let it hold (one 1024-lane workgroup per CU shares
   one box instead of four 256-lane workgroups splitting it). */
static int batch_shape(MCXBatch* b, size_t lblock, int ncu) {
    size_t shmem = batch_params(b, lblock, 0);
    int ldscu = 0, ldsblock = 0;
    const int verbose = getenv("MCXHIP_VERBOSE") != NULL;

    if (hipDeviceGetAttribute(&ldscu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, b->dev) != hipSuccess ||
            hipDeviceGetAttribute(&ldsblock, hipDeviceAttributeMaxSharedMemoryPerBlock, b->dev) != hipSuccess ||
            ldscu <= 0 || ldsblock <= 0 || shmem > (size_t)ldsblock) {
        if (verbose) {
            fprintf(stderr, "mcxhip batch: %zu lanes: LDS per CU %d, per workgroup %d, tables %zu B\n", lblock, ldscu,
                    ldsblock, shmem);
        }

        return 0;
    }

    int nb = mcx_batch_occupancy((int)lblock, shmem, b->reflect);

    if (nb <= 0) {
        if (verbose) {
            fprintf(stderr, "mcxhip batch: %zu lanes: no resident workgroup with %zu B of LDS\n", lblock, shmem);
        }

        return 0;
    }

    size_t per = (size_t)ldscu / (size_t)nb;
    per = per < (size_t)ldsblock ? per : (size_t)ldsblock;
    per = per < MCX_BATCH_LDS_LIMIT ? per : MCX_BATCH_LDS_LIMIT;
    shmem = batch_params(b, lblock, per > shmem + MCX_BATCH_LDS_SLACK ? per - shmem - MCX_BATCH_LDS_SLACK : 0);
    nb = mcx_batch_occupancy((int)lblock, shmem, b->reflect);

    if (nb <= 0 || (size_t)nb * ncu * lblock > b->maxlanes) {
        if (verbose) {
            fprintf(stderr, "mcxhip batch: %zu lanes: %d resident workgroups with %zu B of LDS (lane cap %zu)\n",
                    lblock, nb, shmem, b->maxlanes);
        }

        return 0;
    }

    b->lblock = lblock;
    b->nblock = (size_t)nb * ncu;
    b->shmem = shmem;
    return 1;
}

static double wall_ms_batch(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

static void batch_create_impl(MCXBatch* b, Config** cfgs, unsigned int n, MCXBatchOpt* opt) {
    if (n == 0 || !cfgs) {
        MCX_ERROR(-4, "a batch needs at least one configuration");
    }

    b->n = n;
    b->opt = opt;
    b->cfg = (Config**)host_calloc(n, sizeof(Config*));

    for (unsigned int i = 0; i < n; i++) {
        Config* cfg = cfgs[i];
        b->cfg[i] = cfg;
        mcx_check_supported(cfg);

        if (cfg->mathmode != 0) {
            MCX_ERROR(-4, "the batched launch runs the fast arithmetic only (mathmode 0)");
        }

        if (cfg->issavedet || cfg->seed == SEED_FROM_FILE ||
                (cfg->debuglevel & (MCX_DEBUG_RNG | MCX_DEBUG_MOVE | MCX_DEBUG_MOVE_ONLY)) || cfg->polmedianum ||
                cfg->muamus || cfg->mediabyte > 4 || cfg->omega > 0.f || MCX_IS_ADJOINT_TYPE(cfg->outputtype)) {
            MCX_ERROR(-4, "the batched launch takes no detected photons, replay, RNG dump, trajectories, RF, "
                      "per-voxel or packed media or adjoint outputs");
        }

        if (cfg->srctype == MCX_SRC_PATTERN || cfg->srctype == MCX_SRC_PATTERN3D || cfg->nangle) {
            MCX_ERROR(-4, "the batched launch takes no pattern sources or angle-CDF launches");
        }

        if (cfg->nphoton >= 0xFFFFFFFFull) {
            MCX_ERROR(-4, "more than 2^32-1 photons in one instance");
        }
    }

    if (opt->blocksize && (opt->blocksize < 64 || opt->blocksize > MCX_BATCH_MAXBLOCK || (opt->blocksize & 63))) {
        MCX_ERROR(-4, "blocksize must be a multiple of 64 from 64 to 1024");
    }

    unsigned int activedev = 0;
    int devlist[MAX_DEVICE];
    GPUInfo* gpu = NULL;
    int ndev = mcx_list_gpu(b->cfg[0], &activedev, devlist, &gpu);

    if (ndev == 0 || activedev == 0) {
        free(gpu);
        MCX_ERROR(-99, "No HIP device found (or the specified GPU does not exist)");
    }

    b->dev = devlist[0];
    const int ncu = gpu[b->dev].sm;
    free(gpu);
    BHIP(hipSetDevice(b->dev));
    BHIP(hipStreamCreate(&b->stream));
    BHIP(hipEventCreate(&b->ev0));
    BHIP(hipEventCreate(&b->ev1));
    BHIP(hipEventCreate(&b->ev2));
    BHIP(hipEventCreate(&b->ev3));

    b->prm = (MCXKParam*)host_calloc(n, sizeof(MCXKParam));
    b->norm = (MCXBatchNorm*)host_calloc(n, sizeof(MCXBatchNorm));
    b->fieldoff = (size_t*)host_calloc(n + 1, sizeof(size_t));
    b->dmedia = (uint32_t**)host_calloc(n, sizeof(uint32_t*));
    b->dprop = (float4**)host_calloc(n, sizeof(float4*));
    b->dinvcdf = (float**)host_calloc(n, sizeof(float*));
    b->dsrcdata = (float4**)host_calloc(n, sizeof(float4*));
    b->qof = (uint32_t*)host_calloc(n, sizeof(uint32_t));
    b->order = (uint32_t*)host_calloc(n, sizeof(uint32_t));
    b->weight = (double*)host_calloc(n, sizeof(double));
    /* lists: a multiple of the 8 XCDs, so list q always runs on XCD q % 8; 64
       lists (16 workgroups each at full occupancy) measured best on 256 64^3
       instances, against 2.5x slower with one list and 10% slower with one
       list per instance */
    b->nqueue = opt->nqueue ? opt->nqueue : (n >= 64 ? 64 : ((n + 7) / 8) * 8);

    /* sizes: grids and per-instance input tables, each kind in one allocation */
    size_t nvoxall = 0, npropall = 0, nphaseall = 0, nsrcall = 0, maxlen = 0;

    for (unsigned int i = 0; i < n; i++) {
        Config* cfg = b->cfg[i];
        MCXKParam g;
        int reflect;
        (void)mcx_kparam_setup(cfg, &g, 64, 64, &reflect, 0);
        b->fieldoff[i + 1] = b->fieldoff[i] + g.fieldlen;
        maxlen = g.fieldlen > maxlen ? g.fieldlen : maxlen;
        nvoxall += g.nvox;
        npropall += cfg->medianum;
        nphaseall += (cfg->nphase + 3) & ~3u;
        nsrcall += (size_t)cfg->extrasrclen * 4;
    }

    b->nfield = b->fieldoff[n];
    b->maxlen = maxlen;
    uint32_t* media = batch_alloc<uint32_t>(b, nvoxall);
    float4* prop = batch_alloc<float4>(b, npropall);
    float* phase = batch_alloc<float>(b, nphaseall);
    float4* src = batch_alloc<float4>(b, nsrcall);
    b->dfield = batch_alloc<double>(b, b->nfield);
    b->dout = batch_alloc<float>(b, b->nfield);

    for (unsigned int i = 0; i < n; i++) {
        Config* cfg = b->cfg[i];
        size_t nvox = (size_t)cfg->dim.x * cfg->dim.y * cfg->dim.z;
        b->dmedia[i] = media;
        b->dprop[i] = prop;
        b->dinvcdf[i] = cfg->nphase ? phase : NULL;
        b->dsrcdata[i] = cfg->extrasrclen ? src : NULL;
        BHIP(hipMemcpyAsync(media, cfg->vol, nvox * sizeof(uint32_t), hipMemcpyHostToDevice, b->stream));
        BHIP(hipMemcpyAsync(prop, cfg->prop, cfg->medianum * sizeof(float4), hipMemcpyHostToDevice, b->stream));

        if (cfg->nphase) {
            BHIP(hipMemcpyAsync(phase, cfg->invcdf, cfg->nphase * sizeof(float), hipMemcpyHostToDevice, b->stream));
        }

        if (cfg->extrasrclen) {
            BHIP(hipMemcpyAsync(src, cfg->srcdata, cfg->extrasrclen * sizeof(ExtraSrc), hipMemcpyHostToDevice,
                                b->stream));
        }

        media += nvox;
        prop += cfg->medianum;
        phase += (cfg->nphase + 3) & ~3u;
        src += (size_t)cfg->extrasrclen * 4;

        MCXBatchNorm& nm = b->norm[i];
        nm.field = b->dfield + b->fieldoff[i];
        nm.out = b->dout + b->fieldoff[i];
        nm.len = b->fieldoff[i + 1] - b->fieldoff[i];
        nm.unitinmm = cfg->unitinmm;
        nm.vvox = cfg->steps.x * cfg->steps.y * cfg->steps.z;
        nm.tstep = cfg->tstep;
        nm.outputtype = (uint32_t)cfg->outputtype;
        nm.normalize = cfg->isnormalized ? 1u : 0u;
    }

    b->qlist = (uint32_t*)host_calloc(n, sizeof(uint32_t));
    b->nstate = 2 * b->nqueue + 2 * n;
    b->state0 = (uint32_t*)host_calloc(b->nstate, sizeof(uint32_t));
    b->maxlanes = (size_t)ncu * 2048;   /* 32 waves of 64 lanes per CU at most */
    b->dprm = batch_alloc<MCXKParam>(b, n);
    b->dnorm = batch_alloc<MCXBatchNorm>(b, n);
    b->dqlist = batch_alloc<uint32_t>(b, n);
    b->dstate = batch_alloc<uint32_t>(b, b->nstate);
    b->dstate0 = batch_alloc<uint32_t>(b, b->nstate);
    b->denergy = batch_alloc<double>(b, 2 * (size_t)n);
    b->dcost = batch_alloc<unsigned long long>(b, n);
    b->dnseed = batch_alloc<uint32_t>(b, b->maxlanes * 4);
    BHIP(hipMemcpyAsync(b->dnorm, b->norm, n * sizeof(MCXBatchNorm), hipMemcpyHostToDevice, b->stream));

    /* per-lane RNG seeds from the first instance's seed (src/mcx_host.cpp:696-700) */
    {
        unsigned int* hseed = (unsigned int*)host_calloc(b->maxlanes * 4, sizeof(unsigned int));
        Config* c0 = b->cfg[0];
        mcx_host_seed_words(c0->seed > 0 ? (unsigned int)c0->seed : (unsigned int)time(0), b->maxlanes * 4, hseed);
        hipError_t e = hipMemcpy(b->dnseed, hseed, b->maxlanes * 4 * sizeof(unsigned int), hipMemcpyHostToDevice);
        free(hseed);
        BHIP(e);
    }

    /* Schedule.  The launch shape is the one that keeps the most lanes resident
       (occupancy from the kernel's registers and the LDS its tables need; ties
       go to the larger workgroup) unless the caller fixes it.  Photon costs
       differ by orders of magnitude between instances (scattering events per
       photon scale with mus/mua), so the per-XCD lists are balanced by costs
       measured in a probe step at a scaled-down photon count (the kernel sums
       each instance's workgroup wall-clock time), unless opt->noprobe is set
       (the lists are then balanced by photon counts). */
    /* equal resident lanes go to the first shape listed.  256 lanes comes first:
       on 256 instances of 64^3 with one gate (deposits spread over the whole
       volume) one 1024-lane workgroup per CU with a 4x larger box took 2385 ms
       per step against 2300 ms for four 256-lane workgroups, whose smaller boxes
       are flushed faster at each change of instance */
    const size_t blocks[5] = {256, 512, 1024, 128, 64};
    size_t total = 0;

    for (unsigned int i = 0; i < n; i++) {
        total += b->cfg[i]->nphoton;
        b->weight[i] = (double)b->cfg[i]->nphoton;
    }

    const double t0 = wall_ms_batch();
    size_t bestblock = opt->blocksize;

    if (!bestblock) {
        size_t bestlanes = 0;

        for (int bi = 0; bi < 5; bi++) {
            if (batch_shape(b, blocks[bi], ncu) && b->nblock * b->lblock > bestlanes) {
                bestlanes = b->nblock * b->lblock;
                bestblock = blocks[bi];
            }
        }
    }

    if (!bestblock || !batch_shape(b, bestblock, ncu)) {
        MCX_ERROR(-4, "no launch shape of the batched kernel fits these instances");
    }

    batch_assign(b);

    if (!opt->noprobe) {
        unsigned long long* hcost = (unsigned long long*)host_calloc(n, sizeof(unsigned long long));
        double frac = total > MCX_BATCH_TUNE_PHOTONS ? MCX_BATCH_TUNE_PHOTONS / (double)total : 1.0;

        for (int pass = 0; pass < 2; pass++) {
            size_t np = batch_state(b, frac);
            batch_upload_schedule(b);
            batch_enqueue(b, 1, 1);
            float ms = 0.f;
            BHIP(hipEventSynchronize(b->ev2));
            BHIP(hipEventElapsedTime(&ms, b->ev1, b->ev2));

            /* a short first probe sizes the photon fraction so the next takes ~MCX_BATCH_TUNE_MS */
            if (pass == 0 && frac < 1.0 && ms > 0.f && ms < 0.5 * MCX_BATCH_TUNE_MS) {
                double f = frac * MCX_BATCH_TUNE_MS / ms;
                frac = f < 1.0 ? f : 1.0;
                continue;
            }

            BHIP(hipMemcpy(hcost, b->dcost, n * sizeof(unsigned long long), hipMemcpyDeviceToHost));
            const uint32_t* pool = b->state0 + 2 * b->nqueue;

            for (unsigned int i = 0; i < n; i++) {
                /* ticks per photon times the full photon count */
                b->weight[i] = pool[2 * i + 1] ? (double)hcost[i] / pool[2 * i + 1] * b->cfg[i]->nphoton : 0.0;
            }

            if (getenv("MCXHIP_VERBOSE")) {
                fprintf(stderr, "mcxhip batch tuning: cost probe over %zu photons in %.1f ms\n", np, ms);
            }

            break;
        }

        free(hcost);
        batch_assign(b);
    }

    opt->tunems = (float)(wall_ms_batch() - t0);
    batch_state(b, 1.0);
    batch_upload_schedule(b);
    opt->blocksize = (unsigned int)b->lblock;
    opt->nblock = (unsigned int)b->nblock;
    opt->nqueue = b->nqueue;

    if (getenv("MCXHIP_VERBOSE")) {
        fprintf(stderr, "mcxhip batch: %u instances in %u lists, %zu workgroups of %zu lanes, LDS %zu B, tuning "
                "%.1f ms\n", n, b->nqueue, b->nblock, b->lblock, b->shmem, opt->tunems);
    }

    if (opt->usegraph) {
        BHIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        batch_enqueue(b, 0, 0);
        BHIP(hipStreamEndCapture(b->stream, &b->graph));
        BHIP(hipGraphInstantiate(&b->exec, b->graph, NULL, NULL, 0));
    }
}

static void batch_run_impl(MCXBatch* b, unsigned int nstep) {
    for (unsigned int s = 0; s < nstep; s++) {
        const int last = (s + 1 == nstep);

        if (last) {
            BHIP(hipEventRecord(b->ev0, b->stream));
        }

        if (b->exec) {
            BHIP(hipGraphLaunch(b->exec, b->stream));
        } else {
            batch_enqueue(b, last, 0);
        }

        if (last) {
            BHIP(hipEventRecord(b->ev3, b->stream));
        }
    }

    BHIP(hipStreamSynchronize(b->stream));

    if (nstep) {
        float ms = 0.f;
        BHIP(hipEventElapsedTime(&ms, b->ev0, b->ev3));
        b->opt->stepms = ms;

        if (!b->exec) {
            BHIP(hipEventElapsedTime(&ms, b->ev1, b->ev2));
            b->opt->kernelms = ms;
        }
    }
}

static void batch_collect_impl(MCXBatch* b) {
    double* energy = (double*)host_calloc(2 * (size_t)b->n, sizeof(double));
    float* out = (float*)malloc((b->nfield ? b->nfield : 1) * sizeof(float));

    if (!out) {
        free(energy);
        MCX_ERROR(-4, "out of host memory");
    }

    hipError_t e = hipMemcpy(energy, b->denergy, 2 * (size_t)b->n * sizeof(double), hipMemcpyDeviceToHost);

    if (e == hipSuccess) {
        e = hipMemcpy(out, b->dout, b->nfield * sizeof(float), hipMemcpyDeviceToHost);
    }

    for (unsigned int i = 0; i < b->n && e == hipSuccess; i++) {
        Config* cfg = b->cfg[i];
        size_t len = b->fieldoff[i + 1] - b->fieldoff[i];
        cfg->energyesc = energy[2 * i];
        cfg->energytot = energy[2 * i + 1];
        cfg->energyabs = cfg->energytot - cfg->energyesc;
        cfg->kernelms = b->opt->kernelms;
        cfg->jobms = b->opt->stepms;
        cfg->runtimems = b->opt->stepms;
        cfg->runtime = (unsigned int)(b->opt->stepms + 0.5f);
        const MCXBatchNorm& nm = b->norm[i];

        /* the same float scale as the device finalize */
        if (nm.normalize && cfg->energytot > 0.0) {
            float scale = 1.f;

            if (nm.outputtype == otFlux || nm.outputtype == otFluence) {
                scale = nm.unitinmm / (float)(cfg->energytot * (double)nm.vvox * (double)nm.tstep);

                if (nm.outputtype == otFluence) {
                    scale *= nm.tstep;
                }
            } else if (nm.outputtype == otEnergy || nm.outputtype == otL) {
                scale = 1.f / (float)cfg->energytot;
            }

            cfg->normalizer = scale;
            cfg->his.normalizer = scale;
        }

        if (cfg->issave2pt) {
            if (!cfg->exportfield) {
                cfg->exportfield = (float*)calloc(len * 2, sizeof(float));
            }

            if (cfg->exportfield) {
                memcpy(cfg->exportfield, out + b->fieldoff[i], len * sizeof(float));
            }

            if (cfg->exportfield64) {
                e = hipMemcpy(cfg->exportfield64, b->dfield + b->fieldoff[i], len * sizeof(double),
                              hipMemcpyDeviceToHost);
            }
        }
    }

    free(energy);
    free(out);
    BHIP(e);
}

/* ---------------- C entry points: errors return an id, never exit ---------------- */
extern "C" int mcx_batch_create(Config** cfgs, unsigned int n, MCXBatchOpt* opt, MCXBatch** out) {
    MCXBatch* volatile b = (MCXBatch*)calloc(1, sizeof(MCXBatch));
    *out = NULL;

    if (!b) {
        return -4;
    }

    jmp_buf jb;
    void* old = mcx_push_errjmp(&jb);
    int rc = setjmp(jb);

    if (rc == 0) {
        batch_create_impl(b, cfgs, n, opt);
    }

    mcx_pop_errjmp(old);

    if (rc) {
        batch_free(b);
        return rc;
    }

    *out = b;
    return 0;
}

extern "C" int mcx_batch_run(MCXBatch* b, unsigned int nstep) {
    jmp_buf jb;
    void* old = mcx_push_errjmp(&jb);
    int rc = setjmp(jb);

    if (rc == 0) {
        if (!b) {
            MCX_ERROR(-4, "no batch");
        }

        BHIP(hipSetDevice(b->dev));
        batch_run_impl(b, nstep);
    }

    mcx_pop_errjmp(old);
    return rc;
}

extern "C" int mcx_batch_collect(MCXBatch* b) {
    jmp_buf jb;
    void* old = mcx_push_errjmp(&jb);
    int rc = setjmp(jb);

    if (rc == 0) {
        if (!b) {
            MCX_ERROR(-4, "no batch");
        }

        BHIP(hipSetDevice(b->dev));
        batch_collect_impl(b);
    }

    mcx_pop_errjmp(old);
    return rc;
}

extern "C" void mcx_batch_destroy(MCXBatch* b) {
    if (b) {
        (void)hipSetDevice(b->dev);
        batch_free(b);
    }
}

extern "C" int mcx_run_batch(Config** cfgs, unsigned int n, MCXBatchOpt* opt) {
    MCXBatch* b = NULL;
    int rc = mcx_batch_create(cfgs, n, opt, &b);

    if (rc == 0) {
        rc = mcx_batch_run(b, 1);
    }

    if (rc == 0) {
        rc = mcx_batch_collect(b);
    }

    mcx_batch_destroy(b);
    return rc;
}

/*
 * mcx_oracle.c - CPU oracle for the photon-transport hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the gfx950 kernel
 * is compared against; only tests/, the smoke check and bench.py's CPU
 * baseline load it.  The product (mcxcl_amd) never links or calls it.
 *
 * It is a plain, scalar, single-precision restatement of the reference
 * kernel's photon life cycle, written against the reference's OpenCL
 * branches (src/mcx_core.cl, not the __NVCC__ ones) and its host driver
 * (src/mcx_host.cpp:438-1465):
 *
 *   oracle_run            <- mcx_run_simulation: thread map, glibc seeding,
 *                            time-gate loop, read-back, normalisation
 *   run_thread            <- mcx_main_loop            (src/mcx_core.cl:2307-3307)
 *   launch_photon         <- launchnewphoton          (src/mcx_core.cl:1466-2275)
 *   hitgrid               <- hitgrid                  (src/mcx_core.cl:975-995)
 *   rotatevector          <- rotatevector             (src/mcx_core.cl:1025-1042)
 *   rotatevector2d        <- rotatevector2d           (src/mcx_core.cl:1008-1023)
 *   transmit, reflectcoeff<- transmit/reflectcoeff    (src/mcx_core.cl:1044-1075)
 *   skipvoid              <- skipvoid                 (src/mcx_core.cl:1350-1455)
 *   save_detected         <- finddetector/savedetphoton (src/mcx_core.cl:838-926)
 *   rng_next              <- xorshift128p_nextf       (src/mcx_core.cl:684-698)
 *   media_word_prop       <- updateproperty, packed media formats (src/mcx_core.cl:1079-1193)
 *   pattern_weights       <- the pattern/pattern3d pixel weights (src/mcx_core.cl:1656-1766)
 *   save_traj             <- savedebugdata            (src/mcx_core.cl:929-944)
 *   (in launch_photon)    <- saveexitppath            (src/mcx_core.cl:852-865)
 *   update_stokes         <- rotsphi + updatestokes   (src/mcx_core.cl:781-836)
 *   scatter_angles        <- the polarised (theta, phi) rejection (src/mcx_core.cl:2454-2468)
 *   svmc_update           <- updateproperty_svmc      (src/mcx_core.cl:1231-1278)
 *   ray_plane             <- ray_plane_intersect      (src/mcx_core.cl:1280-1299)
 *   reflect_svmc          <- reflectray_svmc          (src/mcx_core.cl:1302-1345)
 *   (in run_thread)       <- RF forward weights and deposits (src/mcx_core.cl:2750-2760, :2833-2839),
 *                            RF/time-weighted replay deposits (:2568-2623, :2845-2858),
 *                            split-voxel hooks (:2638-2701, :2931-2966, :3069-3243)
 *
 * Every function is pinned by a test that a plausible restatement mistake
 * fails (closed forms, distributions, identities between runs, libm bounds,
 * the reference's published absorbed fractions): tests/test_oracle_pins.py,
 * test_sampler_pins.py, test_output_pins.py, test_replay.py,
 * test_replay_outputs.py, test_media_formats.py, test_trajectory.py,
 * test_rf_forward.py, test_svmc.py, test_exit_tally.py, test_debug_modes.py,
 * test_polarization.py, test_respin.py, test_multisource.py, test_anchors.py.
 *
 * Arithmetic: every transcendental goes through the deterministic FP32
 * functions of mcx_fpmath.h (shared with the kernel's parity mode and pinned
 * against libm in tests/test_oracle_pins.py); division and square root are IEEE.
 * The file is compiled with -ffp-contract=off and without fast-math, so
 * each float operation rounds exactly once, in source order.
 *
 * Tallies: the fluence grid is accumulated in double precision (the "FP64
 * truth" tally); per-thread escaped/launched energies are float, as in the
 * reference (src/mcx_core.cl:2403-2404, :3301-3302), summed in double on the
 * host in thread order (src/mcx_host.cpp:1303-1306).
 *
 * Deliberate deviations from the reference (also applied by the kernel); each
 * removes an out-of-bounds access or a value attributed to the wrong photon,
 * gate or slot by the reference, and changes nothing else:
 *  - the deposit gate index is clamped to maxgate-1 (the reference can write
 *    one gate past the end when (t-twin0)*Rtstep rounds up to maxgate);
 *  - skipvoid returns "not found" when its 3-step entry search ends outside
 *    the volume (the reference then reads media[] outside the volume);
 *  - per-medium detector tallies skip label-0 voxels (the reference indexes
 *    ppath[-1] there); the diffuse-reflectance deposit skips negative gates;
 *  - replay photons are indexed idx*threadphoton+min(idx,oddphoton)+k, the
 *    same map as forward photons; the reference uses min(idx,oddphoton-1)
 *    (src/mcx_core.cl:1591), which replays one photon twice and reads before
 *    the seed buffer when oddphoton is 0;
 *  - after a reflection the boundary code is cleared from isdet (its detector
 *    bit is kept).  The reference keeps it, so the next step that stays in the
 *    voxel restores mediaid = label | code (src/mcx_core.cl:2796, :2928), e.g.
 *    1|3 = 3 after a mirror face, and reads past the property table;
 *  - a photon whose time is NaN leaves the time window (the test is
 *    !(t <= tend) instead of t > tend), so no lane can loop forever;
 *  - scattering-site replay deposits clamp the gate before adding the detector
 *    and source offsets (the reference clamps the sum, src/mcx_core.cl:2591);
 *  - the RF replay phase is the replayed photon's own omega*tof (the reference
 *    reads the next photon's, src/mcx_core.cl:2257-2263), and the imaginary
 *    part of the RFmus deposit goes to the imaginary half (the reference writes
 *    it past its buffer, :2598);
 *  - packed and split-voxel media: the specular entry loss of a source outside
 *    the volume uses the decoded n of the entry voxel (the reference indexes
 *    the label table with the packed word, :1424-1428).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <xmmintrin.h>
#include <pmmintrin.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "mcx_utils.h"
#include "mcx_fpmath.h"

#ifdef ORACLE_FAST
/* CPU-baseline build of the same source (oracle/Makefile, libmcxoracle_fast.so):
   libm single-precision functions instead of the deterministic ones, no event
   counters and float32 private tallies, compiled with -O3 and contraction.  It
   is timed beside the GPU and never used as the checker. */
#define pm_exp(x) expf(x)
#define pm_log(x) logf(x)
#define pm_sin(x) sinf(x)
#define pm_cos(x) cosf(x)
#define pm_acos(x) acosf(x)
#define pm_sincos(x, s, c) (*(s) = sinf(x), *(c) = cosf(x))
#define pm_rsqrt(x) (1.f / sqrtf(x))
#define OCOUNT(w, k) ((void)0)
typedef float OTally;
#else
#define OCOUNT(w, k) ((w)->cnt.k += 1)
typedef double OTally;
#endif

#define K_EPS          FLT_EPSILON
#define R_PI           0.318309886183791f
#define NANGLES        MCX_NANGLES
#define JUST_BELOW_ONE 0.9998f
#define ROULETTE_SIZE  10.f
#define MAX_ACCUM      1000.f

typedef struct {
    float x, y, z, w;
} F4;

typedef struct {
    short x, y, z, w;
} S4;

/* event counters used by the roofline model (algorithmic work per photon) */
typedef struct {
    double nstep;      /* main-loop iterations */
    double nscat;      /* scattering events (direction changes) */
    double ndep;       /* voxel-exit deposits (atomic adds) */
    double nbnd;       /* Fresnel evaluations at refractive-index changes */
    double nlaunch;    /* successful photon launches */
    double nrand;      /* random numbers drawn */
} OracleCounters;

typedef struct {
    F4 srcpos, srcdir, srcparam1, srcparam2;
    F4 maxidx;
    uint32_t dimx, dimxy, nvox, fieldlen;
    float twin0, twin1, oneoverc0, Rtstep, minenergy, minaccumtime;
    uint32_t maxdetphoton, maxmedia, detnum, voidtime, srctype, maxvoidstep, issaveseed, issaveref, isspecular;
    uint32_t maxgate, outputtype, threadphoton, savedetflag, reclen, partialdata, w0offset, mediaformat, gscatter;
    uint32_t is2d, srcnum, doreflect, save2pt, savedet, nphase, debuglevel, extrasrclen;
    uint32_t adjsrc;   /* adjoint outputs: sources 1..adjsrc are the user's, the rest are detectors */
    int32_t seed, oddphoton, replaydet, srcid;
    int reflectblock;  /* the reference compiles its reflection block under MCX_DO_REFLECTION */
    unsigned char bc[12];
    const F4* prop;
    const F4* detpos;
    const uint32_t* media;
    const float* invcdf;
    const ExtraSrc* srcdata;
    const float* replayweight;
    const float* replaytof;
    const int* replaydetid;
    const uint64_t* replayseed;
    const float* muamus;  /* optional per-voxel {mua, mus} in 1/grid unit (2 per voxel) */
    const float* srcpattern;   /* pattern/pattern3d weights, srcnum per pixel (source index fastest) */
    const float* angleinvcdf;  /* launch polar angle inverse CDF in units of pi (nangle entries) */
    uint32_t nangle;
    uint32_t npattern;         /* patterns sharing each photon (srcnum of a pattern source, else 1) */
    uint32_t rfmul;            /* 2 for complex outputs, RF replay or forward (imaginary half after the real) */
    uint32_t rfforward;        /* complex photon weights: omega > 0 without replay */
    float omega;               /* RF replay: angular modulation frequency (rad/s) */
    uint32_t savetraj, maxjumpdebug;  /* trajectory records (src/mcx_core.cl:929-944) */
    uint32_t svmc;                    /* split-voxel media: a second block of nvox plane words */
    uint32_t moveonly;                /* move-only trajectories (debug flag T) */
    uint32_t maxpolmedia;             /* polarised media 1..maxpolmedia (Mie tables in smatrix) */
    uint32_t istrajstokes, trajlen;   /* Stokes vectors in trajectory records; floats per record */
    F4 s0;                            /* launch Stokes vector */
    const F4* smatrix;                /* {S11, S12, S33, S34} x NANGLES per polarised medium */
    uint32_t* detcount;               /* detected photons over all threads (fill-max mode) */
    float* debugdata;                 /* maxjumpdebug records of 6 floats */
    uint32_t* jumpdebug;              /* records written, counted past maxjumpdebug */
} OParam;

typedef struct {
    OTally* field;        /* private field, fieldlen entries (double in the checker) */
    float* det;           /* private detected-photon records */
    uint64_t* seeds;      /* private detected-photon seeds (2 words each) */
    OTally* exitrec;      /* private per-voxel exit tally, issaveref > 1 (maxdetphoton records) */
    float* detw;          /* private exit weights of the detected photons */
    size_t ndet, capdet;
    OracleCounters cnt;
} OWork;

/* ---------------- RNG (src/mcx_core.cl:684-716) ---------------- */
static inline float rng_next(uint64_t t[2], OWork* w) {
    uint64_t s1 = t[0];
    const uint64_t s0 = t[1];
    t[0] = s0;
    s1 ^= s1 << 23;
    t[1] = s1 ^ s0 ^ (s1 >> 18) ^ (s0 >> 5);
    uint32_t lo = (uint32_t)(t[1] + s0);
    OCOUNT(w, nrand);
    return pm_asfloat(0x3F800000u | (lo >> 9)) - 1.0f;
}

static inline float rng_scatlen(uint64_t t[2], OWork* w) {
    return -pm_log(rng_next(t, w) + K_EPS);
}

/* ---------------- geometry helpers ---------------- */
static inline float mcx_nextafterf(float a, int dir) {
    mcx_fbits num;
    num.f = a + 1000.f;
    num.u += (uint32_t)(dir ^ (int)(num.u & 0x80000000u));
    return num.f - 1000.f;
}

static inline float hitgrid(const F4* p, const F4* v, S4* id) {
    float hx = pm_fabs((float)id->x + (float)(v->x > 0.f) - p->x);
    float hy = pm_fabs((float)id->y + (float)(v->y > 0.f) - p->y);
    float hz = pm_fabs((float)id->z + (float)(v->z > 0.f) - p->z);
    hx = pm_fabs((hx + K_EPS) / v->x);
    hy = pm_fabs((hy + K_EPS) / v->y);
    hz = pm_fabs((hz + K_EPS) / v->z);
    float dist = fminf(fminf(hx, hy), hz);
    id->w = (short)(dist == hx ? 0 : (dist == hy ? 1 : 2));
    return dist;
}

static inline void rotatevector(F4* v, float stheta, float ctheta, float sphi, float cphi) {
    if (v->z > -1.f + K_EPS && v->z < 1.f - K_EPS) {
        float tmp0 = 1.f - v->z * v->z;
        float tmp1 = stheta * pm_rsqrt(tmp0);
        F4 n;
        n.x = tmp1 * (v->x * v->z * cphi - v->y * sphi) + v->x * ctheta;
        n.y = tmp1 * (v->y * v->z * cphi + v->x * sphi) + v->y * ctheta;
        n.z = -tmp1 * tmp0 * cphi + v->z * ctheta;
        n.w = v->w;
        *v = n;
    } else {
        F4 n = {stheta * cphi, stheta * sphi, (v->z > 0.f) ? ctheta : -ctheta, v->w};
        *v = n;
    }

    v->x *= pm_rsqrt(v->x * v->x + v->y * v->y + v->z * v->z);
    v->y *= pm_rsqrt(v->x * v->x + v->y * v->y + v->z * v->z);
    v->z *= pm_rsqrt(v->x * v->x + v->y * v->y + v->z * v->z);
}

static inline void rotatevector2d(F4* v, float stheta, float ctheta, int is2d) {
    F4 n = *v;

    if (is2d == 1) {
        n.x = 0.f;
        n.y = v->y * ctheta - v->z * stheta;
        n.z = v->y * stheta + v->z * ctheta;
    } else if (is2d == 2) {
        n.x = v->x * ctheta - v->z * stheta;
        n.y = 0.f;
        n.z = v->x * stheta + v->z * ctheta;
    } else if (is2d == 3) {
        n.x = v->x * ctheta - v->y * stheta;
        n.y = v->x * stheta + v->y * ctheta;
        n.z = 0.f;
    }

    *v = n;
    float tmp0 = pm_rsqrt(v->x * v->x + v->y * v->y + v->z * v->z);
    v->x *= tmp0;
    v->y *= tmp0;
    v->z *= tmp0;
}

static inline void transmit(F4* v, float n1, float n2, int flipdir) {
    float tmp0 = n1 / n2;
    v->x *= tmp0;
    v->y *= tmp0;
    v->z *= tmp0;

    if (flipdir == 0) {
        v->x = pm_sqrt(1.f - v->y * v->y - v->z * v->z) * (float)((v->x > 0.f) - (v->x < 0.f));
    } else if (flipdir == 1) {
        v->y = pm_sqrt(1.f - v->x * v->x - v->z * v->z) * (float)((v->y > 0.f) - (v->y < 0.f));
    } else {
        v->z = pm_sqrt(1.f - v->x * v->x - v->y * v->y) * (float)((v->z > 0.f) - (v->z < 0.f));
    }
}

static inline float reflectcoeff(const F4* v, float n1, float n2, int flipdir) {
    float Icos = pm_fabs((flipdir == 0) ? v->x : (flipdir == 1 ? v->y : v->z));
    float tmp0 = n1 * n1;
    float tmp1 = n2 * n2;
    float tmp2 = 1.f - tmp0 / tmp1 * (1.f - Icos * Icos);

    if (tmp2 > 0.f) {
        float Re = tmp0 * Icos * Icos + tmp1 * tmp2;
        tmp2 = pm_sqrt(tmp2);
        float Im = 2.f * n1 * n2 * Icos * tmp2;
        float Rtotal = (Re - Im) / (Re + Im);
        Re = tmp1 * Icos * Icos + tmp0 * tmp2 * tmp2;
        return (Rtotal + (Re - Im) / (Re + Im)) * 0.5f;
    }

    return 1.f;
}

static inline int inside(const OParam* g, short x, short y, short z) {
    return (unsigned short)x < g->maxidx.x && (unsigned short)y < g->maxidx.y && (unsigned short)z < g->maxidx.z;
}

/* binary16 bits (sign clear) to float: the value m x 2^(e-25) for e > 0,
   m x 2^-24 for the subnormals, exactly */
static float half_to_float(uint32_t h) {
    const uint32_t e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;

    if (e == 31) {
        return m ? NAN : INFINITY;
    }

    return e ? ldexpf((float)(m | 0x400u), (int)e - 25) : ldexpf((float)m, -24);
}

/* src/mcx_core.cl:1079-1193 (updateproperty): a packed media word decoded
   against the label table; the launch sets the properties to t[1] first
   (src/mcx_core.cl:2213) and the formats below overwrite only the components
   they carry, so the others are t[1]'s */
static F4 media_word_prop(uint32_t format, const F4* t, uint32_t mediaid) {
    union {
        uint32_t i;
        float f;
        uint16_t s[2];
        uint8_t c[4];
    } w;
    w.i = mediaid & MED_MASK;
    F4 p = t[1];
    const float nw = t[w.i ? 1 : 0].w;

    switch (format) {
        case MEDIA_MUA_FLOAT:             /* :1089-1096 */
            p.x = fabsf(w.f);
            p.w = nw;
            break;

        case MEDIA_AS_F2H:                /* :1099-1123 */
        case MEDIA_AS_HALF:
            p.x = fabsf(half_to_float(w.s[0] & 0x7FFFu));
            p.y = fabsf(half_to_float(w.s[1] & 0x7FFFu));
            p.w = nw;
            break;

        case MEDIA_LABEL_HALF: {          /* :1126-1150 */
            p = t[w.s[0] & 0x3FFF];
            const float r = fabsf(half_to_float(w.s[1] & 0x7FFFu));
            const int k = (w.s[0] & 0xC000) >> 14;
            if (k == 0) p.x = r;
            if (k == 1) p.y = r;
            if (k == 2) p.z = r;
            if (k == 3) p.w = r;
            break;
        }

        case MEDIA_ASGN_BYTE:             /* :1153-1169 */
            p.x = w.c[0] * (1.f / 255.f) * (t[2].x - t[1].x) + t[1].x;
            p.y = w.c[1] * (1.f / 255.f) * (t[2].y - t[1].y) + t[1].y;
            p.z = w.c[2] * (1.f / 255.f) * (t[2].z - t[1].z) + t[1].z;
            p.w = w.c[3] * (1.f / 127.f) * (t[2].w - t[1].w) + t[1].w;
            break;

        case MEDIA_AS_SHORT:              /* :1172-1186 */
            p.x = w.s[0] * (1.f / 65535.f) * (t[2].x - t[1].x) + t[1].x;
            p.y = w.s[1] * (1.f / 65535.f) * (t[2].y - t[1].y) + t[1].y;
            p.w = nw;
            break;
    }

    return p;
}

static inline F4 getprop(const OParam* g, uint32_t mediaid) {
    if (g->mediaformat > 4) {
        return media_word_prop(g->mediaformat, g->prop, mediaid);
    }

    return g->prop[mediaid & MED_MASK];
}

/* the properties of the photon's voxel: with per-voxel media the label's
   mua and mus are replaced by the voxel's own, g and n stay the label's */
static inline F4 getvoxprop(const OParam* g, uint32_t mediaid, uint32_t idx1d) {
    if (g->mediaformat > 4) {
        return media_word_prop(g->mediaformat, g->prop, mediaid);
    }

    F4 p = g->prop[mediaid & MED_MASK];

    if (g->muamus && (mediaid & MED_MASK)) {
        p.x = g->muamus[2 * (size_t)idx1d];
        p.y = g->muamus[2 * (size_t)idx1d + 1];
    }

    return p;
}

/* ---------------- split-voxel media (svmc) ---------------- */
/* MCXsp (src/mcx_core.cl:594-598): the plane n.x = pd of the photon's voxel,
   n turned away from the photon's side, and the packed labels and flags
   (src/mcx_core.cl:577-589) */
typedef struct {
    float nvx, nvy, nvz, pd;
    uint32_t sv;
} OSplit;

#define SV_LOWER(sv)     ((sv) & 0xFFu)
#define SV_UPPER(sv)     (((sv) >> 8) & 0xFFu)
#define SV_ISSPLIT(sv)   (((sv) >> 16) & 1u)
#define SV_ISUPPER(sv)   (((sv) >> 17) & 1u)
#define SV_FLIP_ISUPPER(sv) ((sv) ^= 0x20000u)
#define SV_CURLABEL(sv)  (SV_ISUPPER(sv) ? SV_UPPER(sv) : SV_LOWER(sv))

static inline void sv_flip(OSplit* sp) {
    sp->nvx = -sp->nvx;
    sp->nvy = -sp->nvy;
    sp->nvz = -sp->nvz;
    sp->pd = -sp->pd;
}

/* updateproperty_svmc (src/mcx_core.cl:1231-1278): the properties at position p
   of the voxel idx1d with integer corner fd.  The label word holds bytes
   {py, px, upper, lower} from the lowest up, the plane word {nz, ny, nx, pz};
   a voxel with upper label 0 is not split.  Outside the grid: the background,
   and the split state is left as it was. */
static F4 svmc_update(const OParam* g, uint32_t mediaid, uint32_t idx1d, const F4* p, const S4* fd, OSplit* sp,
                      F4 prop) {
    if (idx1d == OUTSIDE_VOLUME_MIN || idx1d == OUTSIDE_VOLUME_MAX) {
        return g->prop[0];
    }

    unsigned char c[8];
    const uint32_t lo = g->media[idx1d + g->nvox], hi = mediaid & MED_MASK;

    for (int i = 0; i < 4; i++) {
        c[i] = (unsigned char)(lo >> (8 * i));
        c[i + 4] = (unsigned char)(hi >> (8 * i));
    }

    uint32_t svpacked = (uint32_t)c[7] | ((uint32_t)c[6] << 8);

    if (c[6]) {
        const float rpx = c[5] * (1.f / 255.f) + (float)fd->x;
        const float rpy = c[4] * (1.f / 255.f) + (float)fd->y;
        const float rpz = c[3] * (1.f / 255.f) + (float)fd->z;
        sp->nvx = c[2] * (2.f / 255.f) - 1.f;
        sp->nvy = c[1] * (2.f / 255.f) - 1.f;
        sp->nvz = c[0] * (2.f / 255.f) - 1.f;
        const float r = pm_rsqrt(sp->nvx * sp->nvx + sp->nvy * sp->nvy + sp->nvz * sp->nvz);
        sp->nvx = sp->nvx * r;
        sp->nvy = sp->nvy * r;
        sp->nvz = sp->nvz * r;
        sp->pd = rpx * sp->nvx + rpy * sp->nvy + rpz * sp->nvz;

        if (p->x * sp->nvx + p->y * sp->nvy + p->z * sp->nvz > sp->pd) {
            prop = g->prop[SV_UPPER(svpacked)];
            svpacked |= 0x20000u;
            sv_flip(sp);
        } else {
            prop = g->prop[SV_LOWER(svpacked)];
        }

        svpacked |= 0x10000u;
    } else {
        prop = g->prop[c[7]];
        svpacked &= 0xFFFFu;
    }

    sp->sv = svpacked;
    return prop;
}

/* ray_plane_intersect (src/mcx_core.cl:1280-1299): a step of length *len from p0
   that crosses the plane towards the other side ends on it; *slen becomes the
   scattering length that step uses */
static int ray_plane(const OParam* g, const F4* p0, const F4* v, const F4* prop, float* len, float* slen,
                     const OSplit* sp) {
    const float vdotn = v->x * sp->nvx + v->y * sp->nvy + v->z * sp->nvz;

    if (vdotn <= 0.f) {
        return 0;
    }

    const float d0 = p0->x * sp->nvx + p0->y * sp->nvy + p0->z * sp->nvz - sp->pd;
    const float d1 = d0 + (*len) * vdotn;

    if (d0 * d1 > 0.f) {
        return 0;
    }

    const float len0 = ((*len) * d0) / (d0 - d1);
    *len = (len0 > 0.f) ? len0 : *len;
    *slen = (*len) * prop->y * (v->w + 1.f > (float)g->gscatter ? (1.f - prop->z) : 1.f);
    return 1;
}

/* reflectray_svmc (src/mcx_core.cl:1302-1345): Fresnel reflection or
   refraction of direction c at the plane (sp->nv pointing into the side being
   entered), n1 the index of the side being left.  Returns 1 when the photon
   refracts into label 0. */
static int reflect_svmc(const OParam* g, OWork* w, uint64_t t[2], float n1, float c[3], OSplit* sp, F4* prop) {
    const float icos = pm_fabs(c[0] * sp->nvx + c[1] * sp->nvy + c[2] * sp->nvz);
    const float n2 = SV_ISUPPER(sp->sv) ? g->prop[SV_UPPER(sp->sv)].w : g->prop[SV_LOWER(sp->sv)].w;
    const float tmp0 = n1 * n1, tmp1 = n2 * n2;
    float tmp2 = 1.f - tmp0 / tmp1 * (1.f - icos * icos);

    if (tmp2 > 0.f) {
        float re = tmp0 * icos * icos + tmp1 * tmp2;
        tmp2 = pm_sqrt(tmp2);
        const float im = 2.f * n1 * n2 * icos * tmp2;
        float rtotal = (re - im) / (re + im);
        re = tmp1 * icos * icos + tmp0 * tmp2 * tmp2;
        rtotal = (rtotal + (re - im) / (re + im)) * 0.5f;

        if (rng_next(t, w) <= rtotal) {
            c[0] += (-2.f * icos) * sp->nvx;
            c[1] += (-2.f * icos) * sp->nvy;
            c[2] += (-2.f * icos) * sp->nvz;
            SV_FLIP_ISUPPER(sp->sv);
        } else {
            const float ratio = n1 / n2;
            c[0] += (-icos) * sp->nvx;
            c[1] += (-icos) * sp->nvy;
            c[2] += (-icos) * sp->nvz;
            c[0] = tmp2 * sp->nvx + ratio * c[0];
            c[1] = tmp2 * sp->nvy + ratio * c[1];
            c[2] = tmp2 * sp->nvz + ratio * c[2];
            sv_flip(sp);

            if (SV_CURLABEL(sp->sv) == 0) {
                return 1;
            }

            *prop = g->prop[SV_CURLABEL(sp->sv)];
        }
    } else {
        c[0] += (-2.f * icos) * sp->nvx;
        c[1] += (-2.f * icos) * sp->nvy;
        c[2] += (-2.f * icos) * sp->nvz;
        SV_FLIP_ISUPPER(sp->sv);
    }

    const float r = pm_rsqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    c[0] = c[0] * r;
    c[1] = c[1] * r;
    c[2] = c[2] * r;
    return 0;
}

/* src/mcx_core.cl:1350-1455 */
static int skipvoid(const OParam* g, F4* p, F4* v, F4* f, S4* fd, OSplit* sp) {
    int count = 1, idx1d;
    fd->x = (short)floorf(p->x);
    fd->y = (short)floorf(p->y);
    fd->z = (short)floorf(p->z);
    fd->w = -1;

    while (1) {
        if (inside(g, fd->x, fd->y, fd->z)) {
            idx1d = fd->z * (int)g->dimxy + fd->y * (int)g->dimx + fd->x;

            if (g->media[idx1d] & MED_MASK) {
                p->x -= v->x;
                p->y -= v->y;
                p->z -= v->z;
                fd->x = (short)floorf(p->x);
                fd->y = (short)floorf(p->y);
                fd->z = (short)floorf(p->z);
                f->y -= g->minaccumtime;
                idx1d = fd->z * (int)g->dimxy + fd->y * (int)g->dimx + fd->x;
                count = 0;

                while (!inside(g, fd->x, fd->y, fd->z) || !(g->media[idx1d] & MED_MASK)) {
                    float dist = hitgrid(p, v, fd);
                    f->y += g->minaccumtime * dist;
                    p->x = p->x + dist * v->x;
                    p->y = p->y + dist * v->y;
                    p->z = p->z + dist * v->z;

                    if (fd->w == 0) {
                        fd->x += (v->x > 0.f ? 1 : -1);
                    }

                    if (fd->w == 1) {
                        fd->y += (v->y > 0.f ? 1 : -1);
                    }

                    if (fd->w == 2) {
                        fd->z += (v->z > 0.f ? 1 : -1);
                    }

                    idx1d = fd->z * (int)g->dimxy + fd->y * (int)g->dimx + fd->x;

                    if (count++ > 3) {
                        break;
                    }
                }

                f->y = g->voidtime ? f->y : 0.f;

                if (!inside(g, fd->x, fd->y, fd->z)) {
                    return -1;   /* entry search failed: the reference would read outside the volume */
                }

                F4 htime = g->svmc ? svmc_update(g, g->media[idx1d], (uint32_t)idx1d, p, fd, sp, g->prop[0]) :
                           getprop(g, g->media[idx1d]);

                /* the entry voxel's n: a packed word's decoded n (the reference
                   indexes the label table with the word, src/mcx_core.cl:1424-1428) */
                const float nin = (g->mediaformat > 4) ? htime.w : g->prop[g->media[idx1d] & MED_MASK].w;

                if (g->isspecular && htime.w != g->prop[0].w) {
                    p->w *= 1.f - reflectcoeff(v, g->prop[0].w, nin, fd->w);

                    if (p->w > K_EPS) {
                        transmit(v, g->prop[0].w, nin, fd->w);
                    }
                }

                return idx1d;
            }
        }

        if ((p->x < 0.f && v->x <= 0.f) || (p->x >= g->maxidx.x && v->x >= 0.f) || (p->y < 0.f && v->y <= 0.f) ||
                (p->y >= g->maxidx.y && v->y >= 0.f) || (p->z < 0.f && v->z <= 0.f) ||
                (p->z >= g->maxidx.z && v->z >= 0.f)) {
            return -1;
        }

        p->x = p->x + v->x;
        p->y = p->y + v->y;
        p->z = p->z + v->z;
        fd->x = (short)floorf(p->x);
        fd->y = (short)floorf(p->y);
        fd->z = (short)floorf(p->z);
        f->y += g->minaccumtime;

        if ((uint32_t)count++ > g->maxvoidstep) {
            return -1;
        }
    }
}

typedef struct {
    uint64_t t[2];
    uint64_t photonseed[2];
    F4 p, v, f, prop;
    S4 fd;
    uint32_t idx1d, mediaid;
    float w0, Lmove;
    OSplit sp;            /* the split voxel the photon is in */
    F4 stokes;            /* polarised runs: the Stokes vector {I, Q, U, V} */
    /* per-thread detected-photon scratch (src/mcx_host.cpp:486-492 layout):
       [0, partialdata)   nscat (uint bits) / partial path / momentum per medium
       [partialdata]      escaped energy tally
       [partialdata+1]    launched energy tally
       [partialdata+2]    initial weight of the current photon
       [partialdata+3]    source id of the current photon */
    union {
        float f;
        uint32_t u;
    } pp[4 * 256 + 8];
} OPhoton;

/* src/mcx_core.cl:838-926 */
static void save_detected(const OParam* g, OWork* w, OPhoton* s, uint32_t isdet) {
    int detid = 0;

    if (isdet == OUTSIDE_VOLUME_MIN) {
        detid = -1;
    } else {
        for (uint32_t i = 0; i < g->detnum; i++) {
            if ((g->detpos[i].x - s->p.x) * (g->detpos[i].x - s->p.x) +
                    (g->detpos[i].y - s->p.y) * (g->detpos[i].y - s->p.y) +
                    (g->detpos[i].z - s->p.z) * (g->detpos[i].z - s->p.z) < g->detpos[i].w * g->detpos[i].w) {
                detid = (int)i + 1;
                break;
            }
        }
    }

    if (!detid) {
        return;
    }

    if (g->savedet == FILL_MAXDETPHOTON) {
        /* fill-max mode: a slot past the buffer is given back (src/mcx_core.cl:921-923) */
        if (__atomic_fetch_add(g->detcount, 1u, __ATOMIC_RELAXED) >= g->maxdetphoton) {
            __atomic_fetch_sub(g->detcount, 1u, __ATOMIC_RELAXED);
            return;
        }
    }

    if (w->ndet >= w->capdet) {
        w->capdet = w->capdet ? w->capdet * 2 : 1024;
        w->det = (float*)realloc(w->det, w->capdet * g->reclen * sizeof(float));
        w->seeds = (uint64_t*)realloc(w->seeds, w->capdet * 2 * sizeof(uint64_t));
        w->detw = (float*)realloc(w->detw, w->capdet * sizeof(float));
    }

    w->detw[w->ndet] = s->p.w;
    float* rec = w->det + w->ndet * g->reclen;
    w->seeds[w->ndet * 2] = s->photonseed[0];
    w->seeds[w->ndet * 2 + 1] = s->photonseed[1];
    uint32_t k = 0;

    if (SAVE_DETID(g->savedetflag)) {
        if (g->extrasrclen && g->srcid <= 0) {
            detid |= ((int)s->pp[g->w0offset - 1].f) << 16;
        }

        rec[k++] = (float)detid;
    }

    for (uint32_t i = 0; i < g->partialdata; i++) {
        memcpy(rec + k++, &s->pp[i].u, 4);
    }

    if (SAVE_PEXIT(g->savedetflag)) {
        rec[k++] = s->p.x;
        rec[k++] = s->p.y;
        rec[k++] = s->p.z;
    }

    if (SAVE_VEXIT(g->savedetflag)) {
        rec[k++] = s->v.x;
        rec[k++] = s->v.y;
        rec[k++] = s->v.z;
    }

    if (SAVE_W0(g->savedetflag)) {
        rec[k++] = s->pp[g->w0offset - 2].f;
    }

    if (SAVE_IQUV(g->savedetflag)) {    /* src/mcx_core.cl:915-920 */
        rec[k++] = s->stokes.x;
        rec[k++] = s->stokes.y;
        rec[k++] = s->stokes.z;
        rec[k++] = s->stokes.w;
    }

    w->ndet++;
}

static inline void deposit(const OParam* g, OWork* w, uint32_t addr, float val) {
    (void)g;
    w->field[addr] += (OTally)val;
    OCOUNT(w, ndep);
}

/* time gate of a time-of-flight, clamped to the last gate */
static inline int gate_of(const OParam* g, float t) {
    int gate = (int)floorf((t - g->twin0) * g->Rtstep);
    return gate > (int)g->maxgate - 1 ? (int)g->maxgate - 1 : gate;
}

static inline uint32_t photon_index(const OParam* g, int idx, float fw) {
    int odd = g->oddphoton;
    return (uint32_t)(idx * (int)g->threadphoton + (idx < odd ? idx : odd) + (int)fw);
}

/* the photon id of launch and scattering trajectory records (src/mcx_core.cl:2247,
   :2630): min(idx, (idx < oddphoton) * idx), where termination records (:1501)
   use min(idx, oddphoton) as photon_index does */
static inline uint32_t traj_index(const OParam* g, int idx, float fw) {
    return (uint32_t)(idx * (int)g->threadphoton + (idx < g->oddphoton ? idx : 0) + (int)fw);
}

/* src/mcx_core.cl:929-944: {id as bits, x, y, z, w, (int) source id} */
/* savedebugdata / savedebugstokes (src/mcx_core.cl:929-959): records of 6 floats,
   10 with istrajstokes, the Stokes vector appended in polarised runs */
static void save_traj(const OParam* g, const F4* p, uint32_t id, float srcid, const F4* stokes) {
    uint32_t pos;
#pragma omp atomic capture
    pos = (*g->jumpdebug)++;

    if (pos < g->maxjumpdebug) {
        float* r = g->debugdata + (size_t)pos * g->trajlen;
        memcpy(&r[0], &id, sizeof(id));
        r[1] = p->x;
        r[2] = p->y;
        r[3] = p->z;
        r[4] = p->w;
        r[5] = (float)(int)srcid;

        if (g->maxpolmedia > 0 && g->istrajstokes) {
            r[6] = stokes->x;
            r[7] = stokes->y;
            r[8] = stokes->z;
            r[9] = stokes->w;
        }
    }
}

/* the table entry of scattering angle theta: (uint)(theta NANGLES (1/pi - EPS)) */
static inline const F4* mie_entry(const OParam* g, uint32_t mediaid, float theta) {
    return g->smatrix + (size_t)NANGLES * ((mediaid & MED_MASK) - 1) +
           (uint32_t)(theta * (float)NANGLES * (R_PI - K_EPS));
}

/* rotsphi + updatestokes (src/mcx_core.cl:781-836): into the scattering plane,
   through the Mueller matrix of the angle, into the new meridian plane (u the
   old direction, u2 the new one), normalised to I = 1 */
static void update_stokes(const OParam* g, F4* st, float theta, float phi, const F4* u, const F4* u2,
                          uint32_t mediaid) {
    const float costheta = pm_cos(theta);
    float sin2phi, cos2phi;
    pm_sincos(2.f * phi, &sin2phi, &cos2phi);
    F4 s2 = {st->x, st->y * cos2phi + st->z * sin2phi, -st->y * sin2phi + st->z * cos2phi, st->w};
    const F4* m = mie_entry(g, mediaid, theta);
    st->x = m->x * s2.x + m->y * s2.y;
    st->y = m->y * s2.x + m->x * s2.y;
    st->z = m->z * s2.z + m->w * s2.w;
    st->w = -m->w * s2.z + m->z * s2.w;

    float temp = (u2->z > -1.f && u2->z < 1.f) ? pm_rsqrt((1.f - costheta * costheta) * (1.f - u2->z * u2->z)) : 0.f;
    float cosi = (temp == 0.f) ? 0.f : (((phi > ONE_PI && phi < TWO_PI) ? 1.f : -1.f) * (u2->z * costheta - u->z) * temp);
    cosi = fmaxf(-1.f, fminf(cosi, 1.f));
    const float sini = pm_sqrt(1.f - cosi * cosi);
    const float cos22 = 2.f * cosi * cosi - 1.f;
    const float sin22 = 2.f * sini * cosi;

    s2.x = st->x;
    s2.y = st->y * cos22 - st->z * sin22;
    s2.z = st->y * sin22 + st->z * cos22;
    s2.w = st->w;

    temp = 1.f / s2.x;
    st->y = s2.y * temp;
    st->z = s2.z * temp;
    st->w = s2.w * temp;
    st->x = 1.f;
}

/* launch weights of a pattern pixel (src/mcx_core.cl:1678-1727): one pattern
   scales the photon weight; several share the photon (unit weight), each
   pattern's weight kept per photon, the pixel index in the w0 slot as bits */
static void pattern_weights(const OParam* g, OPhoton* s, float* pp, int pix, float srcw) {
    if (g->srcnum <= 1) {
        s->p.w = srcw * g->srcpattern[pix];
        pp[4] = s->p.w;
    } else {
        const uint32_t bits = (uint32_t)pix;
        memcpy(&pp[2], &bits, sizeof(bits));

        for (uint32_t i = 0; i < g->srcnum; i++) {
            pp[4 + i] = g->srcpattern[(size_t)pix * g->srcnum + i];
        }

        s->p.w = 1.f;
    }
}

/* src/mcx_core.cl:1466-2275; returns 0 when a photon was launched, 1 when the
   thread's photon budget is used up, -1 when no launch succeeded */
static int launch_photon(const OParam* g, OWork* w, OPhoton* s, int idx, uint32_t isdet) {
    float* ppf = (float*)s->pp;
    const uint32_t pd = g->partialdata;
    s->w0 = 1.f;
    s->Lmove = -1.f;

    if (g->savedet == FILL_MAXDETPHOTON && __atomic_load_n(g->detcount, __ATOMIC_RELAXED) >= g->maxdetphoton) {
        return 1;   /* fill-max mode: the detected-photon buffer is full (src/mcx_core.cl:1484-1489) */
    }
    F4 lsrcpos = g->srcpos, lsrcdir = g->srcdir, lp1 = g->srcparam1, lp2 = g->srcparam2;

    if (pm_fabs(s->p.w) >= 0.f) {    /* false only for the NaN of the first call */
        s->pp[pd].f += s->p.w;

        if (g->savetraj) {
            save_traj(g, &s->p, photon_index(g, idx, s->f.w), s->pp[g->w0offset - 1].f, &s->stokes);
        }

        if (s->mediaid == 0 && s->idx1d != OUTSIDE_VOLUME_MIN && s->idx1d != OUTSIDE_VOLUME_MAX && g->issaveref > 1 &&
                s->p.w > 0.f && s->idx1d < g->maxdetphoton) {
            /* saveexitppath (src/mcx_core.cl:852-865): the first column adds the exit
               weight and columns 0..maxmedia-1 each medium's partial path times it,
               both from column 0 as in the reference */
            OTally* rec = w->exitrec + (size_t)s->idx1d * g->reclen;
            rec[0] += s->p.w;

            for (uint32_t i = 0; i < g->maxmedia; i++) {
                rec[i] += s->pp[i].f * s->p.w;
            }
        }

        if (s->mediaid == 0 && s->idx1d != OUTSIDE_VOLUME_MIN && s->idx1d != OUTSIDE_VOLUME_MAX && g->issaveref &&
                s->p.w > 0.f && g->issaveref == 1) {
            int tshift = (int)g->maxgate - 1;
            int t2 = (int)floorf((s->f.y - g->twin0) * g->Rtstep);
            tshift = tshift < t2 ? tshift : t2;

            if (g->extrasrclen && g->srcid < 0) {
                tshift += ((int)s->pp[g->w0offset - 1].f - 1) * (int)g->maxgate;
            }

            if (t2 >= 0 && g->npattern > 1) {
                /* patterns sharing the photon: the interleaved slot of each pattern,
                   scaled by its weight (src/mcx_core.cl:1525-1540) */
                for (uint32_t i = 0; i < g->npattern; i++) {
                    if (pm_fabs(s->pp[g->w0offset + i].f) > 0.f) {
                        deposit(g, w, (s->idx1d + (uint32_t)tshift * g->nvox) * g->npattern + i,
                                -(s->p.w * s->pp[g->w0offset + i].f));
                    }
                }
            } else if (t2 >= 0) {
                deposit(g, w, s->idx1d + (uint32_t)tshift * g->nvox, -s->p.w);
            }
        }

        /* a split voxel's label-0 side counts as outside (src/mcx_core.cl:1556-1559) */
        if (g->savedet && (isdet & DET_MASK) == DET_MASK &&
                (s->mediaid == 0 || (g->svmc && SV_CURLABEL(s->sp.sv) == 0)) && g->issaveref < 2) {
            save_detected(g, w, s, isdet);
        }
    }

    if (g->savedet) {
        for (uint32_t i = 0; i < pd; i++) {
            s->pp[i].u = 0;
        }
    }

    if (s->f.w >= (float)(g->threadphoton + (idx < g->oddphoton))) {
        return 1;
    }

    if (g->seed == SEED_FROM_FILE) {
        uint32_t pid = photon_index(g, idx, s->f.w < 0.f ? 0.f : s->f.w);
        s->t[0] = g->replayseed[pid * 2];
        s->t[1] = g->replayseed[pid * 2 + 1];
    }

    s->photonseed[0] = s->t[0];
    s->photonseed[1] = s->t[1];

    int cursid = 1;

    if (g->extrasrclen && g->srcid != 1) {
        int sid;

        if (g->srcid > 1) {
            sid = g->srcid;
        } else {
            s->pp[g->w0offset - 1].f = (float)((int)(rng_next(s->t, w) * JUST_BELOW_ONE * (float)(g->extrasrclen + 1)) + 1);
            sid = (int)s->pp[g->w0offset - 1].f;
        }

        cursid = sid;

        if (sid > 1) {
            const ExtraSrc* es = g->srcdata + (sid - 2);
            memcpy(&lsrcpos, &es->srcpos, 16);
            memcpy(&lsrcdir, &es->srcdir, 16);
            memcpy(&lp1, &es->srcparam1, 16);
            memcpy(&lp2, &es->srcparam2, 16);
        }
    }

    if (g->seed == SEED_FROM_FILE && g->srcid >= 1) {
        rng_next(s->t, w);
    }

    float* pp = ppf + pd;   /* the reference shifts its ppath pointer here (src/mcx_core.cl:1621) */
    F4* p = &s->p;
    F4* v = &s->v;
    F4* f = &s->f;
    S4* fd = &s->fd;

    do {
        *p = lsrcpos;
        *v = lsrcdir;
        f->x = 0.f;
        f->y = 0.f;
        f->z = g->minaccumtime;
        s->idx1d = pm_asuint(lp2.z);
        s->mediaid = pm_asuint(lp2.w);

        if (g->svmc) {
            s->sp.sv = 0;    /* src/mcx_core.cl:1640-1642 */
        }

        if (g->maxpolmedia > 0) {
            s->stokes = g->s0;    /* src/mcx_core.cl:1633-1638 */
        }
        F4 tmp = {lsrcpos.x, lsrcpos.y, lsrcpos.z, 0.f};
        s->prop = tmp;
        int st = (int)g->srctype;
        int positioned = 0;   /* sources that move the launch point recompute idx1d/mediaid */

        if (st == MCX_SRC_PLANAR || st == MCX_SRC_PATTERN || st == MCX_SRC_PATTERN3D || st == MCX_SRC_FOURIER ||
                st == MCX_SRC_PENCILARRAY) {
            float rx = rng_next(s->t, w);
            float ry = rng_next(s->t, w);

            if (st == MCX_SRC_PATTERN3D) {
                /* a box of param1.x x param1.y x param1.z voxels (src/mcx_core.cl:1661-1667) */
                float rz = rng_next(s->t, w);
                p->x = p->x + rx * lp1.x;
                p->y = p->y + ry * lp1.y;
                p->z = p->z + rz * lp1.z;
                int pix = (int)(rz * JUST_BELOW_ONE * lp1.z) * (int)(lp1.y) * (int)(lp1.x) +
                          (int)(ry * JUST_BELOW_ONE * lp1.y) * (int)(lp1.x) + (int)(rx * JUST_BELOW_ONE * lp1.x);
                pattern_weights(g, s, pp, pix, lsrcpos.w);
            } else {
                p->x = p->x + rx * lp1.x + ry * lp2.x;
                p->y = p->y + rx * lp1.y + ry * lp2.y;
                p->z = p->z + rx * lp1.z + ry * lp2.z;
            }

            if (st == MCX_SRC_PATTERN) {
                /* param1.w x param2.w pixels over the parallelogram (src/mcx_core.cl:1678-1693) */
                int pix = (int)(ry * JUST_BELOW_ONE * lp2.w) * (int)(lp1.w) + (int)(rx * JUST_BELOW_ONE * lp1.w);
                pattern_weights(g, s, pp, pix, lsrcpos.w);
            } else if (st == MCX_SRC_FOURIER) {
                p->w = lsrcpos.w * (pm_cos((floorf(lp1.w) * rx + floorf(lp2.w) * ry + lp1.w - floorf(lp1.w)) * TWO_PI) *
                                    (1.f - lp2.w + floorf(lp2.w)) + 1.f) * 0.5f;
            } else if (st == MCX_SRC_PENCILARRAY) {
                p->x = lsrcpos.x + floorf(rx * lp1.w) * lp1.x / (lp1.w - 1.f) + floorf(ry * lp2.w) * lp2.x / (lp2.w - 1.f);
                p->y = lsrcpos.y + floorf(rx * lp1.w) * lp1.y / (lp1.w - 1.f) + floorf(ry * lp2.w) * lp2.y / (lp2.w - 1.f);
                p->z = lsrcpos.z + floorf(rx * lp1.w) * lp1.z / (lp1.w - 1.f) + floorf(ry * lp2.w) * lp2.z / (lp2.w - 1.f);
            }

            positioned = 1;
            s->prop.x = s->prop.x + (lp1.x + lp2.x) * 0.5f;
            s->prop.y = s->prop.y + (lp1.y + lp2.y) * 0.5f;
            s->prop.z = s->prop.z + (lp1.z + lp2.z) * 0.5f;
            s->prop.w = 0.f;
        } else if (st == MCX_SRC_FOURIERX || st == MCX_SRC_FOURIERX2D) {
            /* param1 = one edge (its .w the other edge's length relative to it), the
               other edge = param1.w dir x param1/|param1|; param2 = the pattern
               (src/mcx_core.cl:1776-1812) */
            float rx = rng_next(s->t, w);
            float ry = rng_next(s->t, w);
            F4 v2 = lp1;
            v2.w *= pm_rsqrt(lp1.x * lp1.x + lp1.y * lp1.y + lp1.z * lp1.z);
            v2.x = v2.w * (lsrcdir.y * lp1.z - lsrcdir.z * lp1.y);
            v2.y = v2.w * (lsrcdir.z * lp1.x - lsrcdir.x * lp1.z);
            v2.z = v2.w * (lsrcdir.x * lp1.y - lsrcdir.y * lp1.x);
            p->x = p->x + rx * lp1.x + ry * v2.x;
            p->y = p->y + rx * lp1.y + ry * v2.y;
            p->z = p->z + rx * lp1.z + ry * v2.z;

            if (st == MCX_SRC_FOURIERX2D) {
                p->w = lsrcpos.w * (pm_sin((lp2.x * rx + lp2.z) * TWO_PI) * pm_sin((lp2.y * ry + lp2.w) * TWO_PI) + 1.f) *
                       0.5f;
            } else {
                p->w = lsrcpos.w * (pm_cos((lp2.x * rx + lp2.y * ry + lp2.z) * TWO_PI) * (1.f - lp2.w) + 1.f) * 0.5f;
            }

            positioned = 1;
            s->prop.x = s->prop.x + (lp1.x + v2.x) * 0.5f;
            s->prop.y = s->prop.y + (lp1.y + v2.y) * 0.5f;
            s->prop.z = s->prop.z + (lp1.z + v2.z) * 0.5f;
            s->prop.w = 0.f;
        } else if (st == MCX_SRC_HYPERBOLOID_GAUSSIAN) {
            /* rays on the one-sheet hyperboloid of a Gaussian beam: waist param1.x,
               param1.y/param1.z the skew; the launch voxel is left as set
               (src/mcx_core.cl:1820-1852) */
            float sphi, cphi;
            float r = TWO_PI * rng_next(s->t, w);
            pm_sincos(r, &sphi, &cphi);
            r = pm_sqrt(0.5f * rng_scatlen(s->t, w)) * lp1.x;
            float hx = -lp1.y / lp1.z;
            float hy = pm_rsqrt(r * r + lp1.z * lp1.z);
            F4 q = {r * (cphi - hx * sphi), r * (sphi + hx * cphi), 0.f, p->w};
            F4 d = {-r * sphi * hy, r * cphi * hy, lp1.z * hy, 0.f};

            if (v->z > -1.f + K_EPS && v->z < 1.f - K_EPS) {
                float t = 1.f - v->z * v->z;
                float stheta = pm_sqrt(t);
                t = pm_rsqrt(t);
                cphi = v->x * t;
                sphi = v->y * t;
                F4 q2 = {q.x * cphi * v->z - q.y * sphi, q.x * sphi * v->z + q.y * cphi, -q.x * stheta, q.w};
                F4 v2 = {d.x * cphi * v->z - d.y * sphi + d.z * cphi * stheta,
                         d.x * sphi * v->z + d.y * cphi + d.z * sphi * stheta, -d.x * stheta + d.z * v->z, v->w};
                q = q2;
                *v = v2;
            } else {
                F4 v2 = {d.x, d.y, (v->z > 0.f) ? d.z : -d.z, v->w};
                *v = v2;
            }

            p->x = q.x + lsrcpos.x;
            p->y = q.y + lsrcpos.y;
            p->z = q.z + lsrcpos.z;
            p->w = q.w;
            F4 c = {lsrcpos.x, lsrcpos.y, lsrcpos.z, 0.f};
            s->prop = c;
            s->Lmove = 0.f;
        } else if (st == MCX_SRC_DISK || st == MCX_SRC_GAUSSIAN || st == MCX_SRC_RING) {
            float sphi, cphi, phi, r;

            if (st == MCX_SRC_DISK || st == MCX_SRC_RING) {
                if (lp1.z > 0.f || lp1.w > 0.f) {
                    phi = pm_fabs(lp1.z - lp1.w) * rng_next(s->t, w) + fminf(lp1.z, lp1.w);
                } else {
                    phi = TWO_PI * rng_next(s->t, w);
                }
            } else {
                phi = TWO_PI * rng_next(s->t, w);
            }

            pm_sincos(phi, &sphi, &cphi);

            if (st == MCX_SRC_DISK || st == MCX_SRC_RING) {
                r = pm_sqrt(rng_next(s->t, w) * pm_fabs(lp1.x * lp1.x - lp1.y * lp1.y) + lp1.y * lp1.y);
            } else if (pm_fabs(lsrcdir.w) < 1e-5f || pm_fabs(lp1.y) < 1e-5f) {
                r = pm_sqrt(-0.5f * pm_log(rng_next(s->t, w))) * lp1.x;
            } else {
                r = lp1.x * lp1.x * ONE_PI / lp1.y;
                r = pm_sqrt(-0.5f * pm_log(rng_next(s->t, w)) * (1.f + (lsrcdir.w * lsrcdir.w / (r * r)))) * lp1.x;
            }

            if (v->z > -1.f + K_EPS && v->z < 1.f - K_EPS) {
                float tmp0 = 1.f - v->z * v->z;
                float tmp1 = r * pm_rsqrt(tmp0);
                p->x = p->x + tmp1 * (v->x * v->z * cphi - v->y * sphi);
                p->y = p->y + tmp1 * (v->y * v->z * cphi + v->x * sphi);
                p->z = p->z - tmp1 * tmp0 * cphi;
            } else {
                p->x += r * cphi;
                p->y += r * sphi;
            }

            positioned = 1;
        } else if (st == MCX_SRC_CONE || st == MCX_SRC_ISOTROPIC || st == MCX_SRC_ARCSINE) {
            float ang, stheta, ctheta, sphi, cphi;
            ang = TWO_PI * rng_next(s->t, w);
            pm_sincos(ang, &sphi, &cphi);

            if (st == MCX_SRC_CONE) {
                do {
                    ang = (lp1.y > 0) ? TWO_PI * rng_next(s->t, w) : pm_acos(2.f * rng_next(s->t, w) - 1.f);
                } while (ang > lp1.x);
            } else if (st == MCX_SRC_ISOTROPIC) {
                ang = pm_acos(2.f * rng_next(s->t, w) - 1.f);
            } else {
                ang = ONE_PI * rng_next(s->t, w);
            }

            pm_sincos(ang, &stheta, &ctheta);
            rotatevector(v, stheta, ctheta, sphi, cphi);
            s->Lmove = 0.f;
        } else if (st == MCX_SRC_ZGAUSSIAN) {
            float ang, stheta, ctheta, sphi, cphi;
            ang = TWO_PI * rng_next(s->t, w);
            pm_sincos(ang, &sphi, &cphi);
            float u1 = rng_next(s->t, w);
            float u2 = rng_next(s->t, w);
            ang = pm_sqrt(-2.f * pm_log(u1)) * (1.f - 2.f * u2) * lp1.x;
            pm_sincos(ang, &stheta, &ctheta);
            rotatevector(v, stheta, ctheta, sphi, cphi);
            s->Lmove = 0.f;
        } else if (st == MCX_SRC_LINE || st == MCX_SRC_SLIT) {
            float r_l = rng_next(s->t, w);
            p->x = p->x + r_l * lp1.x;
            p->y = p->y + r_l * lp1.y;
            p->z = p->z + r_l * lp1.z;

            if (st == MCX_SRC_LINE) {
                float sphi_l, cphi_l;
                r_l = pm_rsqrt(lp1.x * lp1.x + lp1.y * lp1.y + lp1.z * lp1.z);

                if (lp2.x > 0.f) {
                    float ax = lp1.x * r_l, ay = lp1.y * r_l, az = lp1.z * r_l;
                    float vdotaxis = v->x * ax + v->y * ay + v->z * az;
                    v->x = v->x - vdotaxis * ax;
                    v->y = v->y - vdotaxis * ay;
                    v->z = v->z - vdotaxis * az;
                    float vnorm = pm_rsqrt(v->x * v->x + v->y * v->y + v->z * v->z);
                    v->x = v->x * vnorm;
                    v->y = v->y * vnorm;
                    v->z = v->z * vnorm;
                    r_l = lp2.x * (2.f * rng_next(s->t, w) - 1.f);
                    pm_sincos(r_l, &sphi_l, &cphi_l);
                    /* rotate_perpendicular_vector(v, axis, sin, cos) (src/mcx_core.cl:997-1006) */
                    float cx = ay * v->z - az * v->y;
                    float cy = az * v->x - ax * v->z;
                    float cz = ax * v->y - ay * v->x;
                    v->x = v->x * cphi_l + cx * sphi_l;
                    v->y = v->y * cphi_l + cy * sphi_l;
                    v->z = v->z * cphi_l + cz * sphi_l;
                } else {
                    v->x = lp1.x * r_l;
                    v->y = lp1.y * r_l;
                    v->z = lp1.z * r_l;
                    r_l = TWO_PI * rng_next(s->t, w);
                    pm_sincos(r_l, &sphi_l, &cphi_l);
                    rotatevector(v, 1.f, 0.f, sphi_l, cphi_l);
                }
            } else if (lp2.x > 0.f || lp2.y > 0.f) {
                float sphi_s, cphi_s;
                r_l = TWO_PI * rng_next(s->t, w);
                pm_sincos(r_l, &sphi_s, &cphi_s);
                r_l = pm_sqrt(2.f * rng_scatlen(s->t, w));
                cphi_s *= lp2.x * r_l;
                sphi_s *= lp2.y * r_l;
                sphi_s *= pm_rsqrt(lp1.x * lp1.x + lp1.y * lp1.y + lp1.z * lp1.z);
                F4 c = {lp1.y * v->z - lp1.z * v->y, lp1.z * v->x - lp1.x * v->z, lp1.x * v->y - lp1.y * v->x, 0.f};
                s->prop = c;
                cphi_s *= pm_rsqrt(c.x * c.x + c.y * c.y + c.z * c.z);
                v->x += cphi_s * c.x + sphi_s * lp1.x;
                v->y += cphi_s * c.y + sphi_s * lp1.y;
                v->z += cphi_s * c.z + sphi_s * lp1.z;
                r_l = pm_rsqrt(v->x * v->x + v->y * v->y + v->z * v->z);
                v->x *= r_l;
                v->y *= r_l;
                v->z *= r_l;
            }

            positioned = 1;
            s->prop.x = lsrcpos.x + lp1.x * 0.5f;
            s->prop.y = lsrcpos.y + lp1.y * 0.5f;
            s->prop.z = lsrcpos.z + lp1.z * 0.5f;
            s->prop.w = 0.f;
            s->Lmove = -1.f;
        }

        if (positioned) {
            s->idx1d = (uint32_t)((int)floorf(p->z) * (int)g->dimxy + (int)floorf(p->y) * (int)g->dimx + (int)floorf(p->x));

            if (p->x < 0.f || p->y < 0.f || p->z < 0.f || p->x >= g->maxidx.x || p->y >= g->maxidx.y ||
                    p->z >= g->maxidx.z) {
                s->mediaid = 0;
            } else {
                s->mediaid = g->media[s->idx1d];
            }
        }

        if (pm_fabs(p->w) <= g->minenergy) {
            continue;   /* a do-while continue re-tests the loop condition directly */
        }

        if (g->nangle) {
            /* sampled launch polar angle (src/mcx_core.cl:2098-2125): the table holds
               theta/pi at nangle equal steps of the CDF, taken as steps
               (srcdir.w > 0) or linearly interpolated; the azimuth is uniform
               about srcdir (used as is when 0 <= srcdir.w < 1.5) */
            float ang, stheta, ctheta, sphi, cphi;

            if (lsrcdir.w > 0.f) {
                ang = fminf(rng_next(s->t, w) * (float)g->nangle, (float)g->nangle - K_EPS);
                cphi = g->angleinvcdf[(int)ang];
            } else {
                ang = fminf(rng_next(s->t, w) * (float)(g->nangle - 1), (float)(g->nangle - 1) - K_EPS);
                sphi = ang - (float)((int)ang);
                uint32_t i0 = ((uint32_t)ang >= g->nangle - 1) ? g->nangle - 1 : (uint32_t)(int)ang;
                uint32_t i1 = ((uint32_t)ang + 1 >= g->nangle - 1) ? g->nangle - 1 : (uint32_t)((int)ang + 1);
                cphi = (1.f - sphi) * g->angleinvcdf[i0] + sphi * g->angleinvcdf[i1];
            }

            cphi *= ONE_PI;
            pm_sincos(cphi, &stheta, &ctheta);
            ang = TWO_PI * rng_next(s->t, w);
            pm_sincos(ang, &sphi, &cphi);

            if (lsrcdir.w < 1.5f && lsrcdir.w >= 0.f) {
                *v = lsrcdir;
            }

            rotatevector(v, stheta, ctheta, sphi, cphi);
        } else if (s->Lmove < 0.f) {
            if (lsrcdir.w != lsrcdir.w) {   /* NaN focal length: isotropic launch */
                float ang, stheta, ctheta, sphi, cphi;
                ang = TWO_PI * rng_next(s->t, w);
                pm_sincos(ang, &sphi, &cphi);
                ang = pm_acos(2.f * rng_next(s->t, w) - 1.f);
                pm_sincos(ang, &stheta, &ctheta);
                rotatevector(v, stheta, ctheta, sphi, cphi);
            } else if (lsrcdir.w < 0.f && isinf(lsrcdir.w)) {   /* -Inf: Lambertian */
                float ang, stheta, ctheta, sphi, cphi;
                ang = TWO_PI * rng_next(s->t, w);
                pm_sincos(ang, &sphi, &cphi);
                stheta = pm_sqrt(rng_next(s->t, w));
                ctheta = pm_sqrt(1.f - stheta * stheta);
                rotatevector(v, stheta, ctheta, sphi, cphi);
            } else if (lsrcdir.w != 0.f) {   /* focused/diverging beam */
                float Rn2 = (float)((lsrcdir.w > 0.f) - (lsrcdir.w < 0.f));
                s->prop.x += lsrcdir.w * v->x;
                s->prop.y += lsrcdir.w * v->y;
                s->prop.z += lsrcdir.w * v->z;
                v->x = Rn2 * (s->prop.x - p->x);
                v->y = Rn2 * (s->prop.y - p->y);
                v->z = Rn2 * (s->prop.z - p->z);
                Rn2 = pm_rsqrt(v->x * v->x + v->y * v->y + v->z * v->z);
                v->x *= Rn2;
                v->y *= Rn2;
                v->z *= Rn2;
            }
        }

        /* adjoint outputs: a detector source launches from a disk of the detector's
           radius (srcparam1.x) across its direction (src/mcx_core.cl:2159-2184) */
        if (g->adjsrc && cursid > (int)g->adjsrc) {
            float phi = TWO_PI * rng_next(s->t, w), sphi, cphi;
            pm_sincos(phi, &sphi, &cphi);
            float r = pm_sqrt(rng_next(s->t, w)) * lp1.x;

            if (v->z > -1.f + K_EPS && v->z < 1.f - K_EPS) {
                float t0 = 1.f - v->z * v->z;
                float t1 = r * pm_rsqrt(t0);
                p->x += t1 * (v->x * v->z * cphi - v->y * sphi);
                p->y += t1 * (v->y * v->z * cphi + v->x * sphi);
                p->z -= t1 * t0 * cphi;
            } else {
                p->x += r * cphi;
                p->y += r * sphi;
            }

            s->idx1d = (uint32_t)((int)floorf(p->z) * (int)g->dimxy + (int)floorf(p->y) * (int)g->dimx + (int)floorf(p->x));

            if (p->x < 0.f || p->y < 0.f || p->z < 0.f || p->x >= g->maxidx.x || p->y >= g->maxidx.y ||
                    p->z >= g->maxidx.z) {
                s->mediaid = 0;
            } else {
                s->mediaid = g->media[s->idx1d];
            }
        }

        if ((s->mediaid & MED_MASK) == 0) {
            int id = skipvoid(g, p, v, f, fd, &s->sp);

            if (id >= 0) {
                s->idx1d = (uint32_t)id;
                s->mediaid = g->media[s->idx1d];
            }
        }

        fd->x = (short)floorf(p->x);
        fd->y = (short)floorf(p->y);
        fd->z = (short)floorf(p->z);
        s->w0 += 1.f;

        if (s->w0 > (float)g->maxvoidstep) {
            return -1;
        }
    } while ((s->mediaid & MED_MASK) == 0 || pm_fabs(p->w) <= g->minenergy);

    f->w += 1.f;
    s->prop = g->svmc ? svmc_update(g, s->mediaid, s->idx1d, p, fd, &s->sp, s->prop) :
              getvoxprop(g, s->mediaid, s->idx1d);

    if (g->savetraj) {
        /* ppath[3] of the shifted pointer: the source-id slot (src/mcx_core.cl:2247) */
        save_traj(g, p, traj_index(g, idx, f->w), pp[3], &s->stokes);
    }

    pp[1] += p->w;
    s->w0 = p->w;
    pp[2] = (g->srcnum > 1) ? pp[2] : p->w;
    v->w = K_EPS;
    s->Lmove = 0.f;

    if ((g->outputtype == otRF || g->outputtype == otRFmus) && g->seed == SEED_FROM_FILE) {
        /* RF replay: the phase omega * tof of this photon (src/mcx_core.cl:2257-2263,
           which reads the next photon's time of flight; here the photon's own) */
        float rs, rc;
        pm_sincos(g->omega * g->replaytof[photon_index(g, idx, f->w) - 1], &rs, &rc);
        pp[4 + g->srcnum] = rc;
        pp[5 + g->srcnum] = rs;
    }

    OCOUNT(w, nlaunch);
    return 0;
}

/* azimuth and polar angle of one scattering event (src/mcx_core.cl:2446-2500):
   the azimuth is uniform (skipped in 2D, where rotatevector2d draws a sign);
   the polar angle comes from the sampled phase function (nphase > 2, linear
   interpolation of the inverse CDF, :2459-2475) or from Henyey-Greenstein
   inversion, isotropic for |g| <= EPS and after gscatter events (the
   similarity rule, :2484-2497) */
static void scatter_angles(const OParam* g, OWork* w, OPhoton* s, float* stheta, float* ctheta, float* sphi,
                           float* cphi, float* theta_out, float* phi_out) {
    float theta, tmp0;
    *cphi = 1.f;
    *sphi = 0.f;

    if (g->maxpolmedia > 0 && !g->is2d) {
        /* polarised media (src/mcx_core.cl:2454-2468): (theta, phi) by rejection
           against S11 I + S12 (Q cos 2phi + U sin 2phi), bounded by the forward entry */
        float i0, ii, sin2phi, cos2phi;

        do {
            theta = pm_acos(2.f * rng_next(s->t, w) - 1.f);
            tmp0 = TWO_PI * rng_next(s->t, w);
            pm_sincos(2.f * tmp0, &sin2phi, &cos2phi);
            const F4* m0 = mie_entry(g, s->mediaid, 0.f);
            i0 = m0->x * s->stokes.x + m0->y * (s->stokes.y * cos2phi + s->stokes.z * sin2phi);
            const F4* mt = mie_entry(g, s->mediaid, theta);
            ii = mt->x * s->stokes.x + mt->y * (s->stokes.y * cos2phi + s->stokes.z * sin2phi);
        } while (rng_next(s->t, w) * i0 >= ii);

        pm_sincos(tmp0, sphi, cphi);
        pm_sincos(theta, stheta, ctheta);
        *theta_out = theta;
        *phi_out = tmp0;
        return;
    }

    if (!g->is2d) {
        tmp0 = TWO_PI * rng_next(s->t, w);
        pm_sincos(tmp0, sphi, cphi);
    }

    if (g->nphase > 2) {
        tmp0 = rng_next(s->t, w) * (float)(g->nphase - 1);
        theta = tmp0 - (float)((int)tmp0);
        uint32_t i0 = (uint32_t)tmp0 >= g->nphase ? g->nphase - 1 : (uint32_t)(int)tmp0;
        uint32_t i1 = (uint32_t)tmp0 + 1 >= g->nphase ? g->nphase - 1 : (uint32_t)((int)tmp0 + 1);
        tmp0 = (1.f - theta) * g->invcdf[i0] + theta * g->invcdf[i1];
        theta = pm_acos(tmp0);
        *stheta = pm_sin(theta);
        *ctheta = tmp0;
    } else {
        tmp0 = (s->v.w > (float)g->gscatter) ? 0.f : s->prop.z;

        if (pm_fabs(tmp0) > K_EPS) {
            tmp0 = (1.f - s->prop.z * s->prop.z) / (1.f - s->prop.z + 2.f * s->prop.z * rng_next(s->t, w));
            tmp0 *= tmp0;
            tmp0 = (1.f + s->prop.z * s->prop.z - tmp0) / (2.f * s->prop.z);
            tmp0 = fmaxf(-1.f, fminf(1.f, tmp0));
            theta = pm_acos(tmp0);
            *stheta = pm_sin(theta);
            *ctheta = tmp0;
        } else {
            theta = pm_acos(2.f * rng_next(s->t, w) - 1.f);
            pm_sincos(theta, stheta, ctheta);
        }
    }

    /* 2D polarised runs update the Stokes vector with this angle and, where the
       azimuth goes, the sampled cosine (src/mcx_core.cl:2550-2565) */
    *theta_out = theta;
    *phi_out = tmp0;
}

static void run_thread(const OParam* g, OWork* w, int idx, const uint32_t* nseed, float* genergy) {
    OPhoton S;
    OPhoton* s = &S;
    memset(s, 0, sizeof(OPhoton));
    const uint32_t pd = g->partialdata;
    F4* p = &s->p;
    F4* v = &s->v;
    F4* f = &s->f;
    S4* fd = &s->fd;
    p->w = NAN;
    s->mediaid = pm_asuint(g->srcparam2.w);
    fd->w = -1;
    uint32_t isdet = 0, mediaidold = 0, idx1dold;
    float pathlen = 0.f, n1;
    const uint32_t nlimit = g->threadphoton + (idx < g->oddphoton);
    /* split voxels (src/mcx_core.cl:2339-2343): test the plane on this step, and
       whether the step ended on it */
    int testint = 0, hitintf = 0;
    s->sp.sv = 0;
    s->sp.nvx = s->sp.nvy = s->sp.nvz = s->sp.pd = 0.f;

    s->pp[pd].f = genergy[idx * 2];
    s->pp[pd + 1].f = genergy[idx * 2 + 1];
    s->t[0] = (uint64_t)nseed[idx * 4] << 32 | nseed[idx * 4 + 1];
    s->t[1] = (uint64_t)nseed[idx * 4 + 2] << 32 | nseed[idx * 4 + 3];

    if (g->debuglevel & MCX_DEBUG_RNG) {
        return;   /* handled by the caller */
    }

    /* RF forward (omega > 0 without replay): the complex photon weight and its
       value at the last voxel exit, reset at every launch (src/mcx_core.cl:2427-2430) */
    float rwre = 0.f, rwim = 0.f, rw0re = 0.f, rw0im = 0.f;
#define RF_RESET()                       \
    do {                                 \
        if (g->rfforward) {              \
            rwre = rw0re = s->p.w;       \
            rwim = rw0im = 0.f;          \
        }                                \
    } while (0)

    if (launch_photon(g, w, s, idx, 0)) {
        return;
    }

    isdet = s->mediaid & DET_MASK;
    s->mediaid &= MED_MASK;
    RF_RESET();

    while (f->w <= (float)nlimit) {
        OCOUNT(w, nstep);

        if (f->x <= 0.f) {
            f->x = rng_scatlen(s->t, w);

            if (v->w != K_EPS) {
                float cphi, sphi, stheta, ctheta, tmp0, theta = 0.f, phi = 0.f;
                scatter_angles(g, w, s, &stheta, &ctheta, &sphi, &cphi, &theta, &phi);

                /* split voxels tally the label of the photon's side (src/mcx_core.cl:2503-2540) */
                const uint32_t tlab = g->svmc ? SV_CURLABEL(s->sp.sv) : (s->mediaid & MED_MASK);

                if (g->savedet && tlab > 0) {
                    if (SAVE_NSCAT(g->savedetflag)) {
                        s->pp[tlab - 1].u++;
                    }

                    if (SAVE_MOM(g->savedetflag)) {
                        s->pp[g->maxmedia * (SAVE_NSCAT(g->savedetflag) + SAVE_PPATH(g->savedetflag)) +
                                          tlab - 1].f += 1.f - ctheta;
                    }
                }

                const F4 olddir = *v;

                if (g->is2d) {
                    rotatevector2d(v, (rng_next(s->t, w) > 0.5f ? stheta : -stheta), ctheta, (int)g->is2d);
                } else {
                    rotatevector(v, stheta, ctheta, sphi, cphi);
                }

                v->w += 1.f;
                OCOUNT(w, nscat);

                if (g->maxpolmedia > 0) {
                    update_stokes(g, &s->stokes, theta, phi, &olddir, v, s->mediaid);   /* :2550-2565 */
                }

                if (g->outputtype == otWP || g->outputtype == otDCS || g->outputtype == otWPTOF ||
                        (g->seed == SEED_FROM_FILE && g->outputtype == otRFmus)) {
                    /* replay deposits at each scattering site (src/mcx_core.cl:2568-2623):
                       w (WP), (1 - cos theta) w (DCS), tof w (WPTOF), w e^{i omega tof} (RFmus) */
                    uint32_t pid = photon_index(g, idx, f->w) - 1;
                    float im = 0.f;
                    tmp0 = (g->outputtype == otDCS) ? (1.f - ctheta) : 1.f;

                    if (g->outputtype == otRFmus) {
                        const float rw = g->replayweight[pid];
                        tmp0 = rw * s->pp[g->w0offset + g->srcnum].f;
                        im = rw * s->pp[g->w0offset + g->srcnum + 1].f;
                    } else if (g->outputtype == otWPTOF) {
                        tmp0 = g->replaytof[pid];
                        tmp0 *= g->replayweight[pid];
                    } else {
                        tmp0 *= g->replayweight[pid];
                    }

                    int tshift = gate_of(g, g->replaytof[pid]) +
                                 ((g->replaydet == -1) ? (((g->replaydetid[pid] & 0xFFFF) - 1) * (int)g->maxgate) : 0);

                    if (g->extrasrclen && g->srcid < 0) {
                        tshift += ((int)s->pp[g->w0offset - 1].f - 1) * ((g->replaydet == -1) ? (int)g->detnum : 1) *
                                  (int)g->maxgate;
                    }
                    deposit(g, w, s->idx1d + (uint32_t)tshift * g->nvox, tmp0);

                    if (g->outputtype == otRFmus) {
                        deposit(g, w, s->idx1d + (uint32_t)tshift * g->nvox + g->fieldlen, im);
                    }
                }

                if (g->savetraj) {    /* src/mcx_core.cl:2625-2633 */
                    save_traj(g, p, traj_index(g, idx, f->w), s->pp[g->w0offset - 1].f, &s->stokes);
                }
            }

            v->w = (float)(int)v->w;

            if (g->svmc) {
                testint = 1;    /* src/mcx_core.cl:2638-2647 */
            }
        }

        n1 = s->prop.w;
        /* a split voxel's side is taken from the position at every step (src/mcx_core.cl:2656-2672) */
        s->prop = g->svmc ? svmc_update(g, s->mediaid, s->idx1d, p, fd, &s->sp, s->prop) :
                  getvoxprop(g, s->mediaid, s->idx1d);
        f->z = hitgrid(p, v, fd);
        float gfac = (v->w + 1.f > (float)g->gscatter ? (1.f - s->prop.z) : 1.f);
        float slen = f->z * s->prop.y * gfac;
        slen = fminf(slen, f->x);
        f->z = slen / (s->prop.y * gfac);

        if (g->svmc) {    /* src/mcx_core.cl:2684-2701 */
            if (SV_ISSPLIT(s->sp.sv) && testint) {
                float tmplen = f->z;
                hitintf = ray_plane(g, p, v, &s->prop, &tmplen, &slen, &s->sp);
                f->z = tmplen;
            } else {
                hitintf = 0;
            }
        }

        pathlen += f->z;
        p->x = p->x + f->z * v->x;
        p->y = p->y + f->z * v->y;
        p->z = p->z + f->z * v->z;

        if (fd->w == 0) {
            fd->x += (slen == f->x || hitintf) ? 0 : (v->x > 0.f ? 1 : -1);
        }

        if (fd->w == 1) {
            fd->y += (slen == f->x || hitintf) ? 0 : (v->y > 0.f ? 1 : -1);
        }

        if (fd->w == 2) {
            fd->z += (slen == f->x || hitintf) ? 0 : (v->z > 0.f ? 1 : -1);
        }

        if (g->rfforward) {
            /* src/mcx_core.cl:2750-2760: w <- w exp(-(mua + i omega n / c0) ds), p.w = |w| */
            const float att = pm_exp(-s->prop.x * f->z);
            float rs, rc;
            pm_sincos(g->omega * s->prop.w * g->oneoverc0 * f->z, &rs, &rc);
            const float tre = att * (rwre * rc + rwim * rs);
            const float tim = att * (-rwre * rs + rwim * rc);
            rwre = tre;
            rwim = tim;
            p->w = pm_sqrt(rwre * rwre + rwim * rwim);
        } else {
            p->w *= pm_exp(-s->prop.x * f->z);
        }
        f->x -= slen;
        f->y += f->z * s->prop.w * g->oneoverc0;

        {
            const uint32_t plab = g->svmc ? SV_CURLABEL(s->sp.sv) : (s->mediaid & MED_MASK);

            if (g->savedet && SAVE_PPATH(g->savedetflag) && plab > 0) {    /* src/mcx_core.cl:2772-2792 */
                s->pp[g->maxmedia * SAVE_NSCAT(g->savedetflag) + plab - 1].f += f->z;
            }
        }

        mediaidold = s->mediaid | isdet;
        idx1dold = s->idx1d;
        s->idx1d = (uint32_t)(fd->z * (int)g->dimxy + fd->y * (int)g->dimx + fd->x);

        if (!inside(g, fd->x, fd->y, fd->z)) {
            s->mediaid = 0;
            s->idx1d = (fd->x < 0 || fd->y < 0 || fd->z < 0) ? OUTSIDE_VOLUME_MIN : OUTSIDE_VOLUME_MAX;
            isdet = g->bc[(s->idx1d == OUTSIDE_VOLUME_MAX) * 3 + fd->w];
            isdet = ((isdet & 0xF) == bcUnknown) ? (g->doreflect ? bcReflect : bcAbsorb) : isdet;
        } else {
            s->mediaid = g->media[s->idx1d];
            isdet = s->mediaid & DET_MASK;
            s->mediaid &= MED_MASK;
        }

        if ((s->idx1d != idx1dold || hitintf) && idx1dold < g->nvox && mediaidold) {
            if (g->save2pt && f->y >= g->twin0 && f->y < g->twin1) {
                float weight = 0.f, weight_im = 0.f;
                int tshift = gate_of(g, f->y);

                if (g->rfforward) {
                    /* src/mcx_core.cl:2833-2839: (w0 - w) / (mua + i omega n / c0) */
                    const float dre = rw0re - rwre, dim = rw0im - rwim;
                    const float aim = g->omega * s->prop.w * g->oneoverc0;
                    const float amag2 = s->prop.x * s->prop.x + aim * aim;
                    weight = (amag2 < K_EPS) ? (rw0re * pathlen) : (dre * s->prop.x + dim * aim) / amag2;
                    weight_im = (amag2 < K_EPS) ? (rw0im * pathlen) : (dim * s->prop.x - dre * aim) / amag2;
                } else if (g->outputtype == otEnergy) {
                    weight = s->w0 - p->w;
                } else if (g->outputtype == otFluence || g->outputtype == otFlux || MCX_IS_ADJOINT_TYPE(g->outputtype)) {
                    weight = (s->prop.x < K_EPS) ? (s->w0 * pathlen) : ((s->w0 - p->w) / s->prop.x);
                } else if (g->seed == SEED_FROM_FILE) {
                    if (g->outputtype == otJacobian || g->outputtype == otRF || g->outputtype == otWLTOF) {
                        /* replay path-length deposits (src/mcx_core.cl:2845-2858): w L (J),
                           -w L e^{i omega tof} (RF), tof w L (WLTOF) */
                        uint32_t pid = photon_index(g, idx, f->w) - 1;
                        weight = g->replayweight[pid] * pathlen;
                        tshift = gate_of(g, g->replaytof[pid]) +
                                 ((g->replaydet == -1) ? (((g->replaydetid[pid] & 0xFFFF) - 1) * (int)g->maxgate) : 0);

                        if (g->outputtype == otRF) {
                            weight_im = -weight * s->pp[g->w0offset + g->srcnum + 1].f;
                            weight = -weight * s->pp[g->w0offset + g->srcnum].f;
                        } else if (g->outputtype == otWLTOF) {
                            weight = weight * g->replaytof[pid];
                        }
                    }
                } else if (g->outputtype == otL) {
                    weight = s->w0 * pathlen;
                }

                if (g->extrasrclen && g->srcid < 0) {
                    tshift += ((int)s->pp[g->w0offset - 1].f - 1) * ((g->replaydet == -1) ? (int)g->detnum : 1) *
                              (int)g->maxgate;
                }

                if (g->npattern > 1) {
                    /* several patterns share the photon: one slot per pattern, interleaved
                       after the voxel (src/mcx_core.cl:2902-2914) */
                    for (uint32_t i = 0; i < g->npattern; i++) {
                        const float pw = s->pp[g->w0offset + i].f;

                        if (pm_fabs(pw) > 0.f && pm_fabs(weight) > 0.f) {
                            deposit(g, w, (idx1dold + (uint32_t)tshift * g->nvox) * g->npattern + i, weight * pw);
                        }
                    }
                } else if (pm_fabs(weight) > 0.f) {
                    deposit(g, w, idx1dold + (uint32_t)tshift * g->nvox, weight);

                    /* the imaginary part, gated by the real one (:2884-2893); a pattern
                       source's deposit has none (:2902-2914) */
                    if (g->rfforward && g->srctype != MCX_SRC_PATTERN && g->srctype != MCX_SRC_PATTERN3D) {
                        deposit(g, w, idx1dold + (uint32_t)tshift * g->nvox + g->fieldlen, weight_im);
                    }
                }

                if (g->outputtype == otRF && g->seed == SEED_FROM_FILE && g->omega > 0.f && pm_fabs(weight_im) > 0.f) {
                    deposit(g, w, idx1dold + (uint32_t)tshift * g->nvox + g->fieldlen, weight_im);
                }
            }

            s->w0 = p->w;

            if (g->rfforward) {
                rw0re = rwre;
                rw0im = rwim;
            }

            pathlen = 0.f;
        } else {
            s->mediaid = mediaidold;
        }

        if (g->svmc) {    /* src/mcx_core.cl:2931-2952 */
            if (s->idx1d != idx1dold) {
                s->prop = svmc_update(g, s->mediaid, s->idx1d, p, fd, &s->sp, s->prop);
                testint = 1;
            } else if (hitintf) {
                sv_flip(&s->sp);
                SV_FLIP_ISUPPER(s->sp.sv);
                testint = 0;
            }
        }

        if ((s->mediaid == 0 && ((isdet & 0xF) == bcAbsorb || (isdet & 0xF) == bcCyclic ||
                                 ((isdet & 0xF) == bcReflect && n1 == g->prop[0].w))) ||
                (g->svmc && (s->idx1d != idx1dold || hitintf) && !SV_ISUPPER(s->sp.sv) && !SV_LOWER(s->sp.sv) &&
                 (!g->doreflect || n1 == g->prop[0].w)) ||
                !(f->y <= g->twin1)) {
            if (isdet == bcCyclic) {
                if (fd->w == 0) {
                    p->x = mcx_nextafterf(rintf((s->idx1d == OUTSIDE_VOLUME_MIN) ? g->maxidx.x : 0.f),
                                          (v->x > 0.f) - (v->x < 0.f));
                    fd->x = (short)floorf(p->x);
                }

                if (fd->w == 1) {
                    p->y = mcx_nextafterf(rintf((s->idx1d == OUTSIDE_VOLUME_MIN) ? g->maxidx.y : 0.f),
                                          (v->y > 0.f) - (v->y < 0.f));
                    fd->y = (short)floorf(p->y);
                }

                if (fd->w == 2) {
                    p->z = mcx_nextafterf(rintf((s->idx1d == OUTSIDE_VOLUME_MIN) ? g->maxidx.z : 0.f),
                                          (v->z > 0.f) - (v->z < 0.f));
                    fd->z = (short)floorf(p->z);
                }

                if (inside(g, fd->x, fd->y, fd->z)) {
                    s->idx1d = (uint32_t)(fd->z * (int)g->dimxy + fd->y * (int)g->dimx + fd->x);
                    s->mediaid = g->media[s->idx1d];
                    isdet = s->mediaid & DET_MASK;
                    s->mediaid &= MED_MASK;
                    continue;
                }
            }

            uint32_t detflag = (((s->idx1d == OUTSIDE_VOLUME_MAX && g->bc[9 + fd->w]) ||
                                 (s->idx1d == OUTSIDE_VOLUME_MIN && g->bc[6 + fd->w])) ? OUTSIDE_VOLUME_MIN :
                                (mediaidold & DET_MASK));

            if (launch_photon(g, w, s, idx, detflag)) {
                break;
            }

            isdet = s->mediaid & DET_MASK;
            s->mediaid &= MED_MASK;
            RF_RESET();

            if (g->svmc) {
                testint = 1;    /* src/mcx_core.cl:3016-3024 */
            }

            continue;
        }

        if (pm_fabs(p->w) < g->minenergy) {
            if (rng_next(s->t, w) * ROULETTE_SIZE <= 1.f) {
                p->w *= ROULETTE_SIZE;

                if (g->rfforward) {    /* src/mcx_core.cl:3033-3040 */
                    rwre *= ROULETTE_SIZE;
                    rwim *= ROULETTE_SIZE;
                    rw0re *= ROULETTE_SIZE;
                    rw0im *= ROULETTE_SIZE;
                }
            } else {
                if (launch_photon(g, w, s, idx, mediaidold & DET_MASK)) {
                    break;
                }

                isdet = s->mediaid & DET_MASK;
                s->mediaid &= MED_MASK;
                RF_RESET();
                continue;
            }
        }

        if (g->reflectblock) {
            int atface = 1;

            if (g->svmc) {
                /* the split-voxel build updates no property here (its updateproperty is
                   empty); at the plane, reflection or refraction between the labels
                   replaces the voxel-face test (src/mcx_core.cl:3069-3140) */
                if (hitintf) {
                    atface = 0;

                    if (g->prop[SV_LOWER(s->sp.sv)].w != g->prop[SV_UPPER(s->sp.sv)].w) {
                        float c0[3] = {v->x, v->y, v->z};
                        sv_flip(&s->sp);

                        if (reflect_svmc(g, w, s->t, n1, c0, &s->sp, &s->prop)) {
                            if (launch_photon(g, w, s, idx, mediaidold & DET_MASK)) {
                                break;
                            }

                            isdet = s->mediaid & DET_MASK;
                            s->mediaid &= MED_MASK;
                            RF_RESET();
                            testint = 1;
                            continue;
                        }

                        v->x = c0[0];
                        v->y = c0[1];
                        v->z = c0[2];
                    } else {
                        s->prop = g->prop[SV_CURLABEL(s->sp.sv)];
                    }
                }
            } else {
                s->prop = getvoxprop(g, s->mediaid, s->idx1d);
            }

            if (atface && ((s->mediaid && g->doreflect) ||
                    (s->mediaid == 0 && (((isdet & 0xF) == bcUnknown && g->doreflect) || ((isdet & 0xF) == bcReflect ||
                                         (isdet & 0xF) == bcMirror)))) &&
                    (((isdet & 0xF) == bcMirror) || n1 != s->prop.w)) {
                float Rtotal = 1.f;
                float cphi, sphi, stheta, ctheta;
                OCOUNT(w, nbnd);

                if (!g->svmc) {
                    s->prop = getvoxprop(g, s->mediaid, s->idx1d);
                }

                float tmp0 = n1 * n1;
                float tmp1 = s->prop.w * s->prop.w;
                cphi = pm_fabs((fd->w == 0) ? v->x : (fd->w == 1 ? v->y : v->z));
                sphi = 1.f - cphi * cphi;
                f->z = 1.f - tmp0 / tmp1 * sphi;

                if (f->z > 0.f && (isdet & 0xF) != bcMirror) {
                    ctheta = tmp0 * cphi * cphi + tmp1 * f->z;
                    stheta = 2.f * n1 * s->prop.w * cphi * pm_sqrt(f->z);
                    Rtotal = (ctheta - stheta) / (ctheta + stheta);
                    ctheta = tmp1 * cphi * cphi + tmp0 * f->z;
                    Rtotal = (Rtotal + (ctheta - stheta) / (ctheta + stheta)) * 0.5f;
                }

                if (Rtotal < 1.f && (!(s->mediaid == 0 && ((isdet & 0xF) == bcMirror))) &&
                        rng_next(s->t, w) > Rtotal) {
                    transmit(v, n1, s->prop.w, fd->w);

                    if (s->mediaid == 0) {
                        uint32_t detflag = (((s->idx1d == OUTSIDE_VOLUME_MAX && g->bc[9 + fd->w]) ||
                                             (s->idx1d == OUTSIDE_VOLUME_MIN && g->bc[6 + fd->w])) ? OUTSIDE_VOLUME_MIN :
                                            (mediaidold & DET_MASK));

                        if (launch_photon(g, w, s, idx, detflag)) {
                            break;
                        }

                        isdet = s->mediaid & DET_MASK;
                        s->mediaid &= MED_MASK;
                        RF_RESET();
                        continue;
                    }
                } else {
                    /* the reference passes (v>0)-0.5f as an int direction, which truncates to 0:
                       the position snaps to the face without a nudge (src/mcx_core.cl:3205-3209) */
                    if (fd->w == 0) {
                        v->x = -v->x;
                        p->x = mcx_nextafterf(rintf(p->x), 0);
                        fd->x = (short)rintf(p->x);
                    } else if (fd->w == 1) {
                        v->y = -v->y;
                        p->y = mcx_nextafterf(rintf(p->y), 0);
                        fd->y = (short)rintf(p->y);
                    } else {
                        v->z = -v->z;
                        p->z = mcx_nextafterf(rintf(p->z), 0);
                        fd->z = (short)rintf(p->z);
                    }

                    s->idx1d = idx1dold;
                    s->mediaid = (g->media[s->idx1d] & MED_MASK);

                    if (g->svmc) {
                        /* the side the reflected photon is on; label 0 ends it (src/mcx_core.cl:3221-3243) */
                        s->prop = svmc_update(g, s->mediaid, s->idx1d, p, fd, &s->sp, s->prop);

                        if (SV_CURLABEL(s->sp.sv) == 0) {
                            if (launch_photon(g, w, s, idx, mediaidold & DET_MASK)) {
                                break;
                            }

                            isdet = s->mediaid & DET_MASK;
                            s->mediaid &= MED_MASK;
                            RF_RESET();
                            continue;
                        }
                    } else {
                        s->prop = getvoxprop(g, s->mediaid, s->idx1d);
                    }

                    n1 = s->prop.w;
                    isdet &= DET_MASK;
                }
            }

            if (g->moveonly && (s->mediaid == 0 || s->idx1d == OUTSIDE_VOLUME_MIN || s->idx1d == OUTSIDE_VOLUME_MAX)) {
                return;   /* move-only trajectories end the thread, energies unwritten (src/mcx_core.cl:3287-3291) */
            }
        }
    }

    genergy[idx * 2] = s->pp[pd].f;
    genergy[idx * 2 + 1] = s->pp[pd + 1].f;
#undef RF_RESET
}

/* ---------------- host driver ---------------- */
typedef struct {
    OracleCounters cnt;
    double* rawfield;      /* unnormalised double field, fieldlen entries (caller frees) */
    size_t fieldlen;
    float* genergy;        /* per-thread float tallies [2*nthread] (caller frees) */
    unsigned int nthread;
    double seconds;        /* wall-clock of the photon loop */
} OracleResult;

static void set_ieee_fpu(void) {
    /* clear flush-to-zero and denormals-are-zero so results match IEEE hardware */
    _MM_SET_FLUSH_ZERO_MODE(_MM_FLUSH_ZERO_OFF);
    _MM_SET_DENORMALS_ZERO_MODE(_MM_DENORMALS_ZERO_OFF);
}

static double now_sec(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

/*
 * Run the simulation described by a preprocessed Config on the CPU.
 * nthread/threadoffset/threadcount select the thread map: the reference's
 * persistent threads 0..nthread-1 are all seeded (glibc srand/rand), and the
 * threads [threadoffset, threadoffset+threadcount) are simulated (a bounded
 * sample for timing); threadcount=0 runs all of them.  nomp is the number of
 * OpenMP workers (0: all cores).
 * Outputs follow mcx_run_simulation: cfg->exportfield (normalised float),
 * cfg->exportdetected/seeddata/detectedcount, cfg->energytot/energyesc/
 * energyabs and cfg->normalizer.  Returns 0 on success.
 */
/* kernel parameters from a validated Config (the MCXParam fill of
   src/mcx_host.cpp:438-700); allocates *props and *muamus (free both) */
static void oracle_setup(Config* cfg, OParam* gp, F4** props_out, float** muamus_out) {
    memset(gp, 0, sizeof(*gp));
    memcpy(&gp->srcpos, &cfg->srcpos, 16);
    memcpy(&gp->srcdir, &cfg->srcdir, 16);
    memcpy(&gp->srcparam1, &cfg->srcparam1, 16);
    memcpy(&gp->srcparam2, &cfg->srcparam2, 16);
    gp->maxidx.x = (float)cfg->dim.x;
    gp->maxidx.y = (float)cfg->dim.y;
    gp->maxidx.z = (float)cfg->dim.z;
    gp->dimx = cfg->dim.x;
    gp->dimxy = cfg->dim.x * cfg->dim.y;
    gp->nvox = gp->dimxy * cfg->dim.z;
    unsigned int maxgate = (unsigned int)((cfg->tend - cfg->tstart) / cfg->tstep + 0.5);
    cfg->maxgate = maxgate;
    gp->maxgate = maxgate;
    gp->npattern = ((cfg->srctype == MCX_SRC_PATTERN || cfg->srctype == MCX_SRC_PATTERN3D) && cfg->srcnum > 1) ?
                   cfg->srcnum : 1;
    unsigned int nsrcslot = (cfg->srcid < 0) ? cfg->extrasrclen + 1 : gp->npattern;
    unsigned int ndetslot = (cfg->seed == SEED_FROM_FILE && cfg->replaydet == -1) ? cfg->detnum : 1;
    size_t fieldlen = (size_t)gp->nvox * maxgate * nsrcslot * ndetslot;
    gp->fieldlen = (uint32_t)fieldlen;
    gp->rfforward = cfg->omega > 0.f && cfg->seed != SEED_FROM_FILE;
    gp->rfmul = ((cfg->seed == SEED_FROM_FILE && (cfg->outputtype == otRF || cfg->outputtype == otRFmus)) ||
                 gp->rfforward) ? 2 : 1;
    gp->srcpattern = cfg->srcpattern;
    gp->nangle = cfg->nangle;
    gp->angleinvcdf = cfg->angleinvcdf;
    gp->omega = cfg->omega;
    float minstep = fminf(fminf(cfg->steps.x, cfg->steps.y), cfg->steps.z);
    gp->twin0 = cfg->tstart;
    gp->twin1 = cfg->tend;
    gp->oneoverc0 = R_C0 * cfg->unitinmm;
    gp->Rtstep = 1.f / cfg->tstep;
    gp->minenergy = cfg->minenergy;
    gp->minaccumtime = minstep * R_C0 * cfg->unitinmm;
    gp->maxdetphoton = cfg->maxdetphoton;
    gp->maxmedia = cfg->medianum - 1;
    gp->detnum = cfg->detnum;
    gp->voidtime = (uint32_t)cfg->voidtime;
    gp->srctype = (uint32_t)cfg->srctype;
    gp->maxvoidstep = (uint32_t)cfg->maxvoidstep;
    gp->issaveseed = cfg->issaveseed > 0;
    gp->issaveref = (uint32_t)cfg->issaveref;
    gp->isspecular = cfg->isspecular > 0;
    gp->seed = cfg->seed;
    gp->outputtype = (uint32_t)cfg->outputtype;
    gp->savedetflag = cfg->savedetflag;
    gp->partialdata = mcx_partialdata(cfg);
    gp->reclen = mcx_detreclen(cfg);
    gp->w0offset = gp->partialdata + 4;
    gp->mediaformat = cfg->mediabyte;
    gp->svmc = (cfg->mediabyte == MEDIA_2LABEL_SPLIT);
    gp->maxpolmedia = cfg->polmedianum;
    gp->istrajstokes = cfg->istrajstokes ? 1u : 0u;
    gp->trajlen = MCX_DEBUG_REC_LEN + (gp->istrajstokes << 2);
    memcpy(&gp->s0, &cfg->srciquv, sizeof(F4));
    gp->smatrix = (const F4*)cfg->smatrix;
    gp->gscatter = cfg->gscatter;
    gp->srcnum = cfg->srcnum;
    gp->doreflect = (uint32_t)cfg->isreflect;
    gp->save2pt = (uint32_t)cfg->issave2pt;
    gp->savedet = (uint32_t)cfg->issavedet;
    gp->nphase = cfg->nphase;
    gp->invcdf = cfg->invcdf;
    gp->debuglevel = cfg->debuglevel;
    gp->extrasrclen = cfg->extrasrclen;
    /* src/mcx_core.cl:2159-2160: adjoint outputs or srcid -2 */
    gp->adjsrc = ((MCX_IS_ADJOINT_TYPE(cfg->outputtype) || cfg->srcid == -2) && cfg->seed != SEED_FROM_FILE &&
                  cfg->detdir && cfg->detnum && cfg->extrasrclen + 1 > cfg->detnum) ?
                 cfg->extrasrclen + 1 - cfg->detnum : 0;
    gp->srcid = cfg->srcid;
    gp->srcdata = cfg->srcdata;
    gp->replaydet = cfg->replaydet;
    gp->replayweight = cfg->replay.weight;
    gp->replaytof = cfg->replay.tof;
    gp->replaydetid = cfg->replay.detid;
    gp->replayseed = (const uint64_t*)cfg->replay.seed;
    memcpy(gp->bc, cfg->bc, 12);

    /* the reflection block exists when reflection is on, when the boundary flags
       differ from "all absorbing" (as C strings), or when a face reflects or
       mirrors (src/mcx_host.cpp:945-956) */
    {
        char bcs[13] = {0}, allabsorb[7] = {bcAbsorb, bcAbsorb, bcAbsorb, bcAbsorb, bcAbsorb, bcAbsorb, 0};
        memcpy(bcs, cfg->bc, 12);
        gp->reflectblock = cfg->isreflect || (strcmp(bcs, allabsorb) && bcs[0] != 0);

        for (int i = 0; i < 6; i++) {
            if (cfg->bc[i] == bcReflect || cfg->bc[i] == bcMirror) {
                gp->reflectblock = 1;
            }
        }
    }

    unsigned int is2d = (cfg->dim.x == 1 ? 1 : (cfg->dim.y == 1 ? 2 : (cfg->dim.z == 1 ? 3 : 0)));

    if (is2d) {
        is2d = is2d * (((cfg->dim.x > 1) + (cfg->dim.y > 1) + (cfg->dim.z > 1)) == 2);
    }

    gp->is2d = is2d;

    F4* props = (F4*)malloc(sizeof(F4) * cfg->medianum);
    memcpy(props, cfg->prop, sizeof(F4) * cfg->medianum);
    gp->prop = props;
    gp->detpos = (const F4*)cfg->detpos;
    gp->media = cfg->vol;
    float* muamus = NULL;

    if (cfg->muamus) {
        size_t nv = (size_t)cfg->dim.x * cfg->dim.y * cfg->dim.z;
        muamus = (float*)malloc(2 * nv * sizeof(float));

        for (size_t i = 0; i < 2 * nv; i++) {
            muamus[i] = cfg->muamus[i] * cfg->unitinmm;   /* 1/mm -> 1/grid unit, as the label table */
        }
    }

    gp->muamus = muamus;

    *props_out = props;
    *muamus_out = muamus;
}

int oracle_run(Config* cfg, unsigned int nthread, unsigned int threadoffset, unsigned int threadcount, int nomp,
               OracleResult* res) {
    memset(res, 0, sizeof(*res));

    if (nthread == 0) {
        return -1;
    }

    OParam g;
    F4* props;
    float* muamus;
    oracle_setup(cfg, &g, &props, &muamus);
    const unsigned int maxgate = g.maxgate;
    const size_t fieldlen = (size_t)g.fieldlen * g.rfmul;

    /* respin: |respin| repetitions of nphoton/|respin| photons, each on the next
       nthread*4 rand() words (src/mcx_host.cpp:1089-1104, :1319-1332) */
    const unsigned int respin = (unsigned int)(cfg->respin < 0 ? -cfg->respin : (cfg->respin ? cfg->respin : 1));
    size_t nphoton = cfg->nphoton / respin;
    g.threadphoton = (uint32_t)(nphoton / nthread);
    g.oddphoton = (int32_t)(nphoton - (size_t)g.threadphoton * nthread);

    /* glibc seeding, one rand() per 32-bit word (src/mcx_host.cpp:696-700, :759-768) */
    uint32_t* nseed = (uint32_t*)malloc(sizeof(uint32_t) * 4 * (size_t)nthread);

    if (cfg->seed > 0) {
        srand((unsigned int)cfg->seed);
    } else {
        srand((unsigned int)time(0));
    }

    for (size_t j = 0; j < (size_t)nthread * 4; j++) {
        nseed[j] = (uint32_t)rand();
    }

    float* genergy = (float*)calloc((size_t)nthread * 2, sizeof(float));
    unsigned int t0 = threadoffset, t1 = threadcount ? threadoffset + threadcount : nthread;

    if (t1 > nthread) {
        t1 = nthread;
    }

    int nworkers = 1;
#ifdef _OPENMP
    nworkers = nomp > 0 ? nomp : omp_get_max_threads();
#endif
    OWork* works = (OWork*)calloc(nworkers, sizeof(OWork));

    for (int i = 0; i < nworkers; i++) {
        works[i].field = (OTally*)calloc(fieldlen, sizeof(OTally));

        if (cfg->issaveref > 1) {
            works[i].exitrec = (OTally*)calloc((size_t)cfg->maxdetphoton * g.reclen, sizeof(OTally));
        }
    }

    /* trajectory records (src/mcx_core.cl:929-944), one buffer over all repetitions */
    uint32_t njump = 0;
    g.savetraj = (cfg->debuglevel & (MCX_DEBUG_MOVE | MCX_DEBUG_MOVE_ONLY)) ? 1u : 0u;
    g.moveonly = (cfg->debuglevel & MCX_DEBUG_MOVE_ONLY) ? 1u : 0u;
    uint32_t ndetall = 0;
    g.detcount = &ndetall;
    g.maxjumpdebug = cfg->maxjumpdebug;
    g.jumpdebug = &njump;
    g.debugdata = g.savetraj ? (float*)calloc((size_t)cfg->maxjumpdebug * g.trajlen + 1, sizeof(float)) : NULL;

    double tic = now_sec();

    for (unsigned int pass = 0; pass < respin; pass++) {
        if (pass > 0) {
            for (size_t j = 0; j < (size_t)nthread * 4; j++) {
                nseed[j] = (uint32_t)rand();
            }
        }

        if (cfg->debuglevel & MCX_DEBUG_RNG) {
            /* RNG dump: thread idx writes its stream to field[idx], field[idx+nthread], ...
               (src/mcx_core.cl:2408-2414) */
            for (unsigned int idx = t0; idx < t1; idx++) {
                uint64_t t[2];
                t[0] = (uint64_t)nseed[idx * 4] << 32 | nseed[idx * 4 + 1];
                t[1] = (uint64_t)nseed[idx * 4 + 2] << 32 | nseed[idx * 4 + 3];

                for (size_t i = idx; i < fieldlen; i += nthread) {
                    works[0].field[i] = rng_next(t, &works[0]);
                }
            }
        } else {
            #pragma omp parallel num_threads(nworkers)
            {
                int wid = 0;
#ifdef _OPENMP
                wid = omp_get_thread_num();
#endif
                set_ieee_fpu();
                OWork* w = works + wid;
                #pragma omp for schedule(dynamic, 64)

                for (long idx = (long)t0; idx < (long)t1; idx++) {
                    if ((uint32_t)idx >= g.threadphoton * nthread + (uint32_t)g.oddphoton) {
                        continue;
                    }

                    run_thread(&g, w, (int)idx, nseed, genergy);
                }
            }
        }
    }

    res->seconds = now_sec() - tic;

    double* raw = (double*)calloc(fieldlen, sizeof(double));
    size_t ndet = 0;

    for (int i = 0; i < nworkers; i++) {
        for (size_t j = 0; j < fieldlen; j++) {
            raw[j] += works[i].field[j];
        }

        ndet += works[i].ndet;
        res->cnt.nstep += works[i].cnt.nstep;
        res->cnt.nscat += works[i].cnt.nscat;
        res->cnt.ndep += works[i].cnt.ndep;
        res->cnt.nbnd += works[i].cnt.nbnd;
        res->cnt.nlaunch += works[i].cnt.nlaunch;
        res->cnt.nrand += works[i].cnt.nrand;
    }

    if (g.savetraj) {
        const uint32_t nj = njump < cfg->maxjumpdebug ? njump : cfg->maxjumpdebug;

        if (nj) {
            cfg->exportdebugdata = (float*)realloc(cfg->exportdebugdata, ((size_t)cfg->debugdatalen + nj) *
                                                   g.trajlen * sizeof(float));
            memcpy(cfg->exportdebugdata + (size_t)cfg->debugdatalen * g.trajlen, g.debugdata,
                   (size_t)nj * g.trajlen * sizeof(float));
            cfg->debugdatalen += nj;
        }

        free(g.debugdata);
    }

    /* detected photons */
    cfg->detectedcount = 0;
    cfg->his.detected = (unsigned int)ndet;
    size_t nsave = ndet < cfg->maxdetphoton ? ndet : cfg->maxdetphoton;

    if (cfg->issavedet && cfg->issaveref > 1) {
        /* one exit record per voxel, summed over the workers in order; the
           detected count is maxdetphoton (src/mcx_core.cl:3304-3306) */
        const size_t nrec = (size_t)cfg->maxdetphoton * g.reclen;
        free(cfg->exportdetected);
        cfg->exportdetected = (float*)malloc((nrec ? nrec : 1) * sizeof(float));

        for (size_t j = 0; j < nrec; j++) {
            double acc = 0.0;

            for (int i = 0; i < nworkers; i++) {
                acc += works[i].exitrec[j];
            }

            cfg->exportdetected[j] = (float)acc;
        }

        cfg->his.detected = cfg->maxdetphoton;
        cfg->detectedcount = cfg->maxdetphoton;
    } else if (cfg->issavedet) {
        free(cfg->exportdetected);
        cfg->exportdetected = (float*)malloc((nsave ? nsave : 1) * g.reclen * sizeof(float));

        if (cfg->issaveseed) {
            free(cfg->seeddata);
            cfg->seeddata = malloc((nsave ? nsave : 1) * 16);
        }

        size_t k = 0;

        for (int i = 0; i < nworkers && k < nsave; i++) {
            size_t take = works[i].ndet < nsave - k ? works[i].ndet : nsave - k;
            memcpy(cfg->exportdetected + k * g.reclen, works[i].det, take * g.reclen * sizeof(float));

            if (cfg->issaveseed) {
                memcpy((char*)cfg->seeddata + k * 16, works[i].seeds, take * 16);
            }

            if (cfg->exportdetweight && take) {
                memcpy(cfg->exportdetweight + k, works[i].detw, take * sizeof(float));
            }

            k += take;
        }

        cfg->detectedcount = (unsigned int)nsave;
    }

    for (int i = 0; i < nworkers; i++) {
        free(works[i].field);
        free(works[i].exitrec);
        free(works[i].det);
        free(works[i].seeds);
        free(works[i].detw);
    }

    free(works);

    /* energy tallies and normalisation (src/mcx_host.cpp:1299-1306, :1382-1465) */
    cfg->energyesc = 0;
    cfg->energytot = 0;

    for (unsigned int i = 0; i < nthread; i++) {
        cfg->energyesc += genergy[i * 2];
        cfg->energytot += genergy[i * 2 + 1];
    }

    cfg->energyabs = cfg->energytot - cfg->energyesc;

    if (cfg->issave2pt) {
        free(cfg->exportfield);
        cfg->exportfield = (float*)calloc(fieldlen, sizeof(float));

        for (size_t i = 0; i < fieldlen; i++) {
            cfg->exportfield[i] = (float)raw[i];
        }

        if (cfg->isnormalized && !(cfg->debuglevel & MCX_DEBUG_RNG)) {
            /* src/mcx_host.cpp:1382-1465 */
            const size_t half = g.fieldlen;   /* the real part; RF outputs have an imaginary half after it */
            float Vvox = cfg->steps.x * cfg->steps.y * cfg->steps.z;
            float scale = 1.f;
            const int isreplayout = cfg->outputtype == otJacobian || cfg->outputtype == otWP ||
                                    cfg->outputtype == otDCS || cfg->outputtype == otRF ||
                                    cfg->outputtype == otRFmus || cfg->outputtype == otWLTOF ||
                                    cfg->outputtype == otWPTOF;

            if (cfg->outputtype == otFlux || cfg->outputtype == otFluence || MCX_IS_ADJOINT_TYPE(cfg->outputtype) ||
                    g.rfforward) {
                scale = cfg->unitinmm / (float)(cfg->energytot * Vvox * cfg->tstep);

                if (cfg->outputtype == otFluence || MCX_IS_ADJOINT_TYPE(cfg->outputtype)) {
                    scale *= cfg->tstep;
                }
            } else if (cfg->outputtype == otEnergy || cfg->outputtype == otL) {
                scale = 1.f / (float)cfg->energytot;
            } else if (isreplayout) {
                if (cfg->seed == SEED_FROM_FILE && cfg->replaydet == -1) {
                    /* per detector (src/mcx_host.cpp:1398-1422), both halves of RF
                       outputs; with srcid < 0 per (source, detector) slice, each by its
                       own photons' weights (the reference scales only the first
                       source's slices) */
                    const int nsrc = (cfg->extrasrclen && cfg->srcid < 0) ? (int)cfg->extrasrclen + 1 : 1;

                    for (int sr = 1; sr <= nsrc; sr++) {
                        for (int d = 1; d <= (int)cfg->detnum; d++) {
                            float sd = 0.f;

                            for (size_t i = 0; i < cfg->nphoton; i++)
                                if ((cfg->replay.detid[i] & 0xFFFF) == d && (nsrc == 1 || (cfg->replay.detid[i] >> 16) == sr)) {
                                    sd += cfg->replay.weight[i];
                                }

                            if (sd > 0.f) {
                                sd = cfg->unitinmm / sd;
                            }

                            for (uint32_t part = 0; part < g.rfmul; part++) {
                                mcx_normalize(cfg->exportfield + part * half +
                                              ((size_t)(sr - 1) * cfg->detnum + (d - 1)) * g.nvox * maxgate, sd,
                                              (int)(g.nvox * maxgate), 1, 0, 1);
                            }
                        }
                    }

                    scale = 0.f;
                } else {
                    float sw = 0.f;

                    for (size_t i = 0; i < cfg->nphoton; i++) {
                        sw += cfg->replay.weight[i];
                    }

                    scale = sw > 0.f ? cfg->unitinmm / sw : 0.f;
                }
            }

            if (cfg->extrasrclen && cfg->srcid < 0) {
                scale *= (float)(cfg->extrasrclen + 1);
            }

            if (scale != 0.f) {
                cfg->normalizer = scale;
                cfg->his.normalizer = scale;

                if (g.npattern > 1) {
                    /* several patterns: pattern i scaled by npixel / (its summed weights)
                       (src/mcx_host.cpp:1351-1364, :1436-1447), Kahan sums in float */
                    const int psize = (cfg->srctype == MCX_SRC_PATTERN3D) ?
                                      (int)cfg->srcparam1.x * (int)cfg->srcparam1.y * (int)cfg->srcparam1.z :
                                      (int)cfg->srcparam1.w * (int)cfg->srcparam2.w;

                    for (uint32_t i = 0; i < g.npattern; i++) {
                        float sum = 0.f, c = 0.f;

                        for (int j = 0; j < psize; j++) {
                            float y = cfg->srcpattern[(size_t)j * g.npattern + i] - c;
                            float t = sum + y;
                            c = (t - sum) - y;
                            sum = t;
                        }

                        mcx_normalize(cfg->exportfield, (float)psize / sum * scale, (int)(fieldlen / g.npattern),
                                      cfg->isnormalized, (int)i, (int)g.npattern);
                    }
                } else {
                    mcx_normalize(cfg->exportfield, scale, (int)fieldlen, cfg->isnormalized, 0, 1);
                }
            }
        }
    }

    if (cfg->exportfield64) {
        memcpy(cfg->exportfield64, raw, fieldlen * sizeof(double));
    }

    res->rawfield = raw;
    res->fieldlen = fieldlen;
    res->genergy = genergy;
    res->nthread = nthread;
    free(nseed);
    free(props);
    free(muamus);
    return 0;
}

void oracle_free_result(OracleResult* res) {
    free(res->rawfield);
    free(res->genergy);
    memset(res, 0, sizeof(*res));
}

/* first n words of glibc rand() after srand(seed) (pins the host seeding) */
void oracle_glibc_rand(unsigned int seed, unsigned int n, uint32_t* out) {
    srand(seed);

    for (unsigned int i = 0; i < n; i++) {
        out[i] = (uint32_t)rand();
    }
}

/* first n floats of the xorshift128+ stream from 4 seed words */
void oracle_rng_stream(const uint32_t seed4[4], unsigned int n, float* out, uint64_t* state_out) {
    OWork w;
    memset(&w, 0, sizeof(w));
    uint64_t t[2];
    t[0] = (uint64_t)seed4[0] << 32 | seed4[1];
    t[1] = (uint64_t)seed4[2] << 32 | seed4[3];

    for (unsigned int i = 0; i < n; i++) {
        out[i] = rng_next(t, &w);
    }

    if (state_out) {
        state_out[0] = t[0];
        state_out[1] = t[1];
    }
}

/* expose the geometric kernels for brute-force checks */
float oracle_hitgrid(const float p[3], const float v[3], const short id[3], int* face) {
    F4 pp = {p[0], p[1], p[2], 0.f}, vv = {v[0], v[1], v[2], 0.f};
    S4 ii = {id[0], id[1], id[2], 0};
    float d = hitgrid(&pp, &vv, &ii);
    *face = ii.w;
    return d;
}

float oracle_reflectcoeff(const float v[3], float n1, float n2, int flipdir) {
    F4 vv = {v[0], v[1], v[2], 0.f};
    return reflectcoeff(&vv, n1, n2, flipdir);
}

void oracle_rotatevector(float v[3], float stheta, float ctheta, float sphi, float cphi) {
    F4 vv = {v[0], v[1], v[2], 0.f};
    rotatevector(&vv, stheta, ctheta, sphi, cphi);
    v[0] = vv.x;
    v[1] = vv.y;
    v[2] = vv.z;
}

/* single-precision math functions, for accuracy checks */
void oracle_fpmath(int which, const float* x, float* y, size_t n) {
    for (size_t i = 0; i < n; i++) {
        switch (which) {
            case 0:
                y[i] = pm_exp(x[i]);
                break;

            case 1:
                y[i] = pm_log(x[i]);
                break;

            case 2:
                y[i] = pm_sin(x[i]);
                break;

            case 3:
                y[i] = pm_cos(x[i]);
                break;

            case 4:
                y[i] = pm_acos(x[i]);
                break;

            case 5:
                y[i] = pm_rsqrt(x[i]);
                break;

            default:
                y[i] = x[i];
        }
    }
}

size_t oracle_result_sizeof(void) {
    return sizeof(OracleResult);
}

/* Launch sampler check: n successive launches of thread 0 (its seed words are
   the first four rand() words after srand(cfg->seed)), each recorded right
   after launch_photon as 8 floats {p.x, p.y, p.z, p.w, v.x, v.y, v.z, label}.
   Returns the number of launches recorded (a launch that fails ends the run). */
int oracle_launch_samples(Config* cfg, unsigned int n, float* out) {
    OParam g;
    F4* props;
    float* muamus;
    oracle_setup(cfg, &g, &props, &muamus);
    g.threadphoton = n;
    g.oddphoton = 0;
    g.savedet = 0;
    g.save2pt = 0;
    set_ieee_fpu();

    OWork w;
    memset(&w, 0, sizeof(w));
    OPhoton* s = (OPhoton*)calloc(1, sizeof(OPhoton));
    srand((unsigned int)(cfg->seed > 0 ? cfg->seed : 1));
    uint32_t words[4];

    for (int i = 0; i < 4; i++) {
        words[i] = (uint32_t)rand();
    }

    s->t[0] = (uint64_t)words[0] << 32 | words[1];
    s->t[1] = (uint64_t)words[2] << 32 | words[3];
    s->p.w = NAN;
    s->mediaid = pm_asuint(g.srcparam2.w);
    unsigned int k = 0;

    for (; k < n; k++) {
        if (launch_photon(&g, &w, s, 0, 0)) {
            break;
        }

        float* o = out + (size_t)k * 8;
        o[0] = s->p.x;
        o[1] = s->p.y;
        o[2] = s->p.z;
        o[3] = s->p.w;
        o[4] = s->v.x;
        o[5] = s->v.y;
        o[6] = s->v.z;
        o[7] = (float)(s->mediaid & MED_MASK);
        s->p.w = -1.f;   /* the next call tallies a zero-weight exit: no deposit, no record */
        s->mediaid = 1;
    }

    free(s);
    free(props);
    free(muamus);
    return (int)k;
}

/* Scattering sampler check: n scattering events of a photon in medium `label`
   that has already scattered `nscat` times (gscatter compares against it),
   travelling along (0.6, 0, 0.8) before each event; 3 floats per event
   {ctheta, stheta, cos of the angle between the old and new directions}. */
void oracle_scatter_samples(Config* cfg, unsigned int label, unsigned int nscat, unsigned int n, float* out) {
    OParam g;
    F4* props;
    float* muamus;
    oracle_setup(cfg, &g, &props, &muamus);
    set_ieee_fpu();
    OWork w;
    memset(&w, 0, sizeof(w));
    OPhoton* s = (OPhoton*)calloc(1, sizeof(OPhoton));
    uint32_t seed4[4] = {0x2545F491u, 0x4F6CDD1Du, (uint32_t)cfg->seed, 0x9E3779B9u};
    s->t[0] = (uint64_t)seed4[0] << 32 | seed4[1];
    s->t[1] = (uint64_t)seed4[2] << 32 | seed4[3];
    s->prop = g.prop[label];

    for (unsigned int k = 0; k < n; k++) {
        float stheta, ctheta, sphi, cphi, theta, phi;
        F4 v0 = {0.6f, 0.f, 0.8f, (float)nscat};
        s->v = v0;
        scatter_angles(&g, &w, s, &stheta, &ctheta, &sphi, &cphi, &theta, &phi);

        if (g.is2d) {
            rotatevector2d(&s->v, (rng_next(s->t, &w) > 0.5f ? stheta : -stheta), ctheta, (int)g.is2d);
        } else {
            rotatevector(&s->v, stheta, ctheta, sphi, cphi);
        }

        out[(size_t)k * 3] = ctheta;
        out[(size_t)k * 3 + 1] = stheta;
        out[(size_t)k * 3 + 2] = s->v.x * v0.x + s->v.y * v0.y + s->v.z * v0.z;
    }

    free(s);
    free(props);
    free(muamus);
}

/* Polarised scattering check: n events in polarised medium `label` of a photon
   travelling along (0.6, 0, 0.8) with Stokes vector stokes4 before each; 9
   floats per event {theta, phi, new direction x, y, z, new I, Q, U, V}. */
void oracle_pol_scatter_samples(Config* cfg, unsigned int label, const float* stokes4, unsigned int n, float* out) {
    OParam g;
    F4* props;
    float* muamus;
    oracle_setup(cfg, &g, &props, &muamus);
    set_ieee_fpu();
    OWork w;
    memset(&w, 0, sizeof(w));
    OPhoton* s = (OPhoton*)calloc(1, sizeof(OPhoton));
    uint32_t seed4[4] = {0x2545F491u, 0x4F6CDD1Du, (uint32_t)cfg->seed, 0x9E3779B9u};
    s->t[0] = (uint64_t)seed4[0] << 32 | seed4[1];
    s->t[1] = (uint64_t)seed4[2] << 32 | seed4[3];
    s->prop = g.prop[label];
    s->mediaid = label;

    for (unsigned int k = 0; k < n; k++) {
        float stheta, ctheta, sphi, cphi, theta = 0.f, phi = 0.f;
        F4 v0 = {0.6f, 0.f, 0.8f, 1.f};
        s->v = v0;
        memcpy(&s->stokes, stokes4, sizeof(F4));
        scatter_angles(&g, &w, s, &stheta, &ctheta, &sphi, &cphi, &theta, &phi);
        rotatevector(&s->v, stheta, ctheta, sphi, cphi);
        update_stokes(&g, &s->stokes, theta, phi, &v0, &s->v, s->mediaid);
        float* o = out + (size_t)k * 9;
        o[0] = theta;
        o[1] = phi;
        o[2] = s->v.x;
        o[3] = s->v.y;
        o[4] = s->v.z;
        o[5] = s->stokes.x;
        o[6] = s->stokes.y;
        o[7] = s->stokes.z;
        o[8] = s->stokes.w;
    }

    free(s);
    free(props);
    free(muamus);
}

/*
 * mcx_core.h - shared declarations between the gfx950 photon-transport kernels
 * (mcx_core_*.hip) and the host runtime (mcx_host.hip).
 */
#ifndef MCX_CORE_H
#define MCX_CORE_H

#include <stdint.h>

#include <hip/hip_runtime.h>

/* Workgroup size bounds of the transport kernel.  The bit-reproducible build
   launches 256-lane workgroups.  The fast build runs one workgroup per CU, so
   that a single LDS deposit box can use most of the CU's LDS instead of being
   split between several workgroups: 1024 lanes (16 waves, <=128 VGPRs) for the
   pencil-beam specialisation (109 VGPRs) and for the general path without
   detected photons or without reflection (7-9 VGPRs spilled to scratch, still
   6-9% faster than 768 lanes), 768 lanes (12 waves, <=168 VGPRs, no spills)
   for the general path with both, which would spill 21. */
#define MCX_MAX_BLOCK 256
#define MCX_FAST_SIMPLE_BLOCK 1024
#define MCX_FAST_BLOCK 1024
#define MCX_FAST_DETREF_BLOCK 768
#define MCX_FAST_BOUND(SAVEDET, REFLECT, SIMPLE) \
    ((SIMPLE) ? MCX_FAST_SIMPLE_BLOCK : (((SAVEDET) && (REFLECT)) ? MCX_FAST_DETREF_BLOCK : MCX_FAST_BLOCK))
/* The full general path (pattern, fourierx, hyperboloid and angle-CDF launches,
   shared patterns, RF and time-weighted replay outputs) is a separate instance so
   the common general path keeps its registers: 768 lanes (<=168 VGPRs). */
#define MCX_FAST_FULL_BLOCK 768
#define MCX_FAST_BOUND_FULL(SAVEDET, REFLECT, SIMPLE, FULL) \
    ((FULL) ? MCX_FAST_FULL_BLOCK : MCX_FAST_BOUND(SAVEDET, REFLECT, SIMPLE))
/* Workgroup size bound of the batched kernel (general path at four waves per
   SIMD, <=128 VGPRs): up to one 1024-lane workgroup per CU. */
#define MCX_BATCH_BLOCK_MAX 1024

/* Kernel parameter block, passed by value (lives in the kernarg segment, read
   with scalar loads).  Field meanings follow the reference's MCXParam
   (src/mcx_host.h:106-166), restricted to what this kernel reads. */
struct MCXKParam {
    float4 srcpos, srcdir, srcparam1, srcparam2;
    float maxidx_x, maxidx_y, maxidx_z;
    uint32_t dimx, dimy, dimz, dimxy, nvox;
    float twin0, twin1, oneoverc0, Rtstep, minenergy, minaccumtime;
    uint32_t maxdetphoton, maxmedia, medianum, detnum, voidtime, srctype, maxvoidstep;
    uint32_t issaveseed, issaveref, isspecular, maxgate, outputtype, threadphoton, savedetflag, reclen;
    uint32_t partialdata, w0offset, gscatter, is2d, srcnum, doreflect, save2pt, savedet, nphase, extrasrclen;
    uint32_t nthread, debugrng, mediaformat;
    uint32_t adjsrc;              /* adjoint outputs: sources 1..adjsrc are the user's, the rest are
                                     detectors launched from a disk (0: no adjoint output) */
    int32_t seed, oddphoton, replaydet, srcid;
    uint32_t fieldlen;
    /* deposit privatisation: the box [x0,x0+nx)x[y0,y0+ny)x[z0,z0+nz) of the first ng
       gates is accumulated in LDS per workgroup (float64 cells in the
       bit-reproducible build, float32 in the fast build) and flushed once; ng=0
       disables.  RF forward runs keep a second set of ng box gates after the first
       for the imaginary half: box gate ng + t flushes to field gate box_imgate + t
       (box_imgate = imoff / nvox; 0 when the box holds no imaginary half) */
    uint32_t box_x0, box_y0, box_z0, box_nx, box_ny, box_nz, box_ng, box_cells;
    uint32_t box_imgate;
    uint32_t stokesoff;           /* polarised runs: the Stokes vector's offset in the lane scratch */
    uint32_t rfoff;               /* RF forward: offset of the weight at the last voxel exit (re, im) */
    uint32_t svoff;               /* split voxels: offset of the current voxel's plane {nx, ny, nz, pd} */
    uint32_t seedoff;             /* saved seeds: offset of the RNG state at launch (4 words) */
    uint32_t nrep;                /* global accumulator replicas (power of two) */
    uint32_t npattern;            /* patterns sharing each photon (a pattern source with srcnum > 1), else 1;
                                     their slots interleave after the voxel: (vox + gate*nvox)*npattern + k */
    uint32_t nangle;              /* entries of the launch polar-angle table (0: none) */
    uint32_t pplen;               /* floats of per-lane LDS scratch (records, tallies, w0, source id,
                                     pattern weights, the RF replay phase) */
    uint32_t imoff;               /* complex outputs (RF replay, RF forward): offset of the imaginary half
                                     (0: real output) */
    float omega;                  /* angular modulation frequency, rad/s */
    uint32_t rfforward;           /* 1: complex photon weights and fluence (omega > 0 without replay) */
    uint32_t dynamic;             /* 1: photons are handed out per workgroup at run time
                                     (threadphoton/oddphoton then split photons over workgroups) */
    uint32_t simple;              /* 1: pencil beam from a non-void voxel without focus, extra sources,
                                     replay, 2D, phase table, reflectance tally or replay output types */
    uint32_t full;                /* 1: the run uses a pattern, fourierx or hyperboloid launch, an
                                     angle-CDF launch, shared patterns, RF/time-weighted replay outputs,
                                     packed media words or trajectory records */
    uint32_t savetraj;            /* 1: photon positions are recorded at launch, scattering and termination */
    uint32_t maxjumpdebug;        /* trajectory records the buffer holds */
    uint32_t svmc;                /* 1: split-voxel media, two words per voxel (the second block of
                                     nvox words holds each voxel's plane) */
    uint32_t moveonly;            /* 1: move-only trajectories: a lane's run ends when its photon moves
                                     into the background without being relaunched */
    uint32_t maxpolmedia;         /* polarised media (1..maxpolmedia carry Mie tables); 0: unpolarised */
    uint32_t istrajstokes;        /* 1: trajectory records carry the Stokes vector (10 floats, else 6) */
    uint32_t trajlen;             /* floats per trajectory record */
    float4 s0;                    /* launch Stokes vector {I, Q, U, V} */
    const float4* smatrix;        /* {S11, S12, S33, S34} at 181 angles (0..180 degrees) per polarised medium */
    const float* polcdf;          /* fast build: the 182 cos(theta) bin edges, then per polarised medium the
                                     182-entry CDF of the angle bins weighted by S11 (theta's marginal) */
    unsigned char bc[12];

    const uint32_t* media;        /* label volume, x fastest; bit 31 marks detector-adjacent voxels */
    const float4* prop;           /* {mua, mus, g, n} per label */
    const float4* detpos;         /* {x, y, z, r} per detector */
    const float* invcdf;          /* sampled phase-function inverse CDF (nphase) */
    const float* srcpattern;      /* pattern/pattern3d pixel weights, npattern per pixel */
    const float* angleinvcdf;     /* launch polar angle / pi at nangle equal CDF steps */
    const float4* srcdata;        /* extra sources, 4 float4 each */
    const float* replayweight;
    const float* replaytof;
    const int* replaydetid;
    const uint64_t* replayseed;   /* 2 words per replayed photon */
    const uint32_t* nseed;        /* 4 words per thread */
    double* field;                /* fluence accumulator, nrep x fieldlen entries */
    const float2* muamus;         /* optional per-voxel {mua, mus} (1/grid unit); NULL: from the label */
    float* detweight;             /* optional exit weight of each detected photon */
    float* genergy;               /* {escaped, launched} per thread */
    float* detrec;                /* detected-photon records, reclen floats each */
    uint64_t* detseed;            /* 2 words per detected photon */
    uint32_t* detcount;
    float* debugdata;             /* trajectory records: {photon id bits, x, y, z, w, source id} */
    uint32_t* jumpdebug;          /* trajectory records written, counted past maxjumpdebug */
};

/* per-instance output of a batched step: grid, float output and normalisation */
struct MCXBatchNorm {
    const double* field;
    float* out;
    size_t len;
    float unitinmm, vvox, tstep;
    uint32_t outputtype, normalize, pad;
};

#ifdef __cplusplus
extern "C" {
#endif

/* launch the transport kernel; variant flags select template instances */
hipError_t mcx_launch_parity(const MCXKParam* prm, int nblock, int nthreadblock, size_t shmem, hipStream_t stream,
                             int savedet, int reflect);
hipError_t mcx_launch_fast(const MCXKParam* prm, int nblock, int nthreadblock, size_t shmem, hipStream_t stream,
                           int savedet, int reflect);
/* RNG stream dump: field[i] for i = idx; i < fieldlen; i += nthread */
/* resident workgroups per CU of the fast transport kernel variant for this launch shape */
/* adjoint outputs from the normalised per-source field (kind 0: phi_s phi_d,
   kind 1: grad phi_s . grad phi_d), scaled, optionally by a per-label factor;
   complex when field_im is given (RF), the imaginary part into out_im */
hipError_t mcx_launch_adjoint(const float* field, const float* field_im, float* out, float* out_im, uint32_t nvox,
                              uint32_t nx, uint32_t ny, uint32_t maxgate, uint32_t ns, uint32_t nd, int kind,
                              float scale, const float* opscale, const uint32_t* media, hipStream_t stream);
int mcx_fast_occupancy(int nthreadblock, size_t shmem, int savedet, int reflect, int simple);
hipError_t mcx_launch_atomic_probe(double* grid, uint32_t n, uint32_t nblock, uint32_t k, hipStream_t s);
hipError_t mcx_launch_rngdump(const MCXKParam* prm, int nblock, int nthreadblock, hipStream_t stream);
/* batched launch: many simulations in one persistent kernel (fast arithmetic) */
hipError_t mcx_launch_batch(const MCXKParam* prm, const uint32_t* qlist, uint32_t nqueue, uint32_t* state,
                            uint32_t* nseed, double* energy, unsigned long long* cost, int nblock, int nthreadblock,
                            size_t shmem, int reflect, hipStream_t stream);
int mcx_batch_occupancy(int nthreadblock, size_t shmem, int reflect);
hipError_t mcx_launch_batch_prep(double* field, size_t nfield, double* energy, uint32_t nenergy, uint32_t* state,
                                 const uint32_t* state0, uint32_t nstate, hipStream_t stream);
hipError_t mcx_launch_batch_finalize(const MCXBatchNorm* norm, const double* energy, uint32_t ninst, size_t maxlen,
                                     hipStream_t stream);
/* sum the nrep replicas into in[0..n), then out[i] = (float)in[i] * scale (out may be NULL) */
hipError_t mcx_launch_finalize(double* in, float* out, size_t n, unsigned int nrep, float scale, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif

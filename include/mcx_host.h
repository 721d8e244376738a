/*
 * mcx_host.h - the simulation entry points of the MI355X build (C ABI).
 *
 * Drop-in for the reference's host runtime header (src/mcx_host.h:46-48,
 * :168-174).  mcx_run_simulation keeps its signature and its contract: all
 * inputs and outputs travel through the Config (exportfield, exportdetected,
 * seeddata, detectedcount, energytot/energyesc/energyabs, normalizer,
 * runtime, his); the two trailing pointers are unused, as in the reference.
 *
 * mcx_list_gpu keeps its name, its GPUInfo output and its activedev count but
 * drops the OpenCL handle types: it returns the number of HIP devices and
 * fills activedevlist with HIP device ordinals (documented ABI change; the
 * reference returns a cl_platform_id and fills cl_device_id handles).
 */
#ifndef MCX_AMD_HOST_H
#define MCX_AMD_HOST_H

#include "mcx_utils.h"

#ifdef __cplusplus
extern "C" {
#endif

/* replaces void mcx_run_simulation(Config*, float*, float*) (src/mcx_host.h:168) */
void mcx_run_simulation(Config* cfg, float* fluence, float* totalenergy);

/* replaces cl_platform_id mcx_list_gpu(Config*, unsigned int*, cl_device_id*, GPUInfo**)
   (src/mcx_host.h:169); returns the number of usable devices */
int mcx_list_gpu(Config* cfg, unsigned int* activedev, int* activedevlist, GPUInfo** info);

/* non-exiting variant of mcx_run_simulation for language bindings: returns 0 on
   success, the error id otherwise (message in mcx_last_error()) */
int mcx_try_run_simulation(Config* cfg);

/* No counterpart in the reference.  For callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc): a forward fluence run of a validated Config (as
   mcx_run_simulation takes it, without detected photons, replay, adjoint
   outputs, RNG dump or per-voxel media) on the CURRENT HIP device, with the
   volume and the outputs in caller-owned device memory:
     media    uint32[nmedia]    cfg->vol as validated (labels, bit 31 detector mask)
     field64  double[fieldlen]  unnormalised accumulator (cleared here)
     fluence  float[fieldlen]   normalised float output, or NULL
   fieldlen = voxels x gates (x sources when srcid < 0); stream is a
   hipStream_t (NULL: the null stream).  The energies and the normaliser are
   returned in cfg; the call returns once the outputs are complete.  Nothing
   is allocated: per-lane seeds and tallies and the staged tables live in a
   6.1 MB per-device static scratch, so a run may use at most 262144 lanes
   (the fast mode's resident wave on 256 CUs; set nthread in bit-reproducible
   mode), 4096 media, 4096 phase samples and 256 extra sources; larger runs
   are rejected with an error.  Calls on one device share that scratch (and
   a host staging chunk): overlapping calls race.  Returns 0, or the error id
   with the message in mcx_last_error(). */
int mcx_run_simulation_noalloc(Config* cfg, const unsigned int* media, size_t nmedia, double* field64, float* fluence,
                               size_t fieldlen, void* stream);

/* number of HIP devices visible to this process (0 when none) */
int mcx_device_count(void);

/* first n words of the seeding generator after seeding with seed: the
   glibc srand()/rand() sequence the reference uses to seed its per-thread
   RNG streams (src/mcx_host.cpp:696-700, :759-768) */
void mcx_host_seed_words(unsigned int seed, size_t n, unsigned int* out);

/* self-checks of the device arithmetic, for the test suite: evaluate the
   kernel's single-precision math (which: 0 exp, 1 log, 2 sin, 3 cos, 4 acos,
   5 rsqrt, 6 divide x[i]/x[i+n], 7 sqrt) in parity mode (mode 1) or fast mode
   (mode 0) on device `gpuid`; returns 0 on success */
int mcx_device_fpmath(int gpuid, int which, int mode, const float* x, float* y, size_t n);

/* first n floats of the device xorshift128+ stream seeded with seed4[4] */
int mcx_device_rng_stream(int gpuid, const unsigned int seed4[4], unsigned int n, float* out);

/* measured throughput (atomics/s) of float64 atomic adds at hashed addresses of
   an ncell-cell grid on device gpuid: the ceiling the fluence deposits run against */
int mcx_device_atomic_ceiling(int gpuid, unsigned int ncell, double* rate);

/* launch-shape model of the fast transport kernel variant (savedet, reflect,
   simple: 1 the pencil-beam path, 0 the general path, 2 the full general path
   with pattern/fourierx/hyperboloid/angle-CDF launches and RF replay outputs) on
   the current device: its VGPRs per lane and largest workgroup
   (returns 0 on success), and the workgroups of nthreadblock lanes with shmem
   bytes of dynamic LDS that one CU holds (gfx950: 4 SIMDs of at most 8 waves
   and 512 VGPRs per lane each, the device's LDS per CU) */
int mcx_fast_kernel_attr(int savedet, int reflect, int simple, int* vgpr, int* maxlanes);
int mcx_fast_occupancy(int nthreadblock, size_t shmem, int savedet, int reflect, int simple);

/* ---- batched launch (no counterpart in the reference, whose host runs one
   Config per call): many independent simulations with their own volumes,
   media, sources and photon counts run in one persistent launch ---- */
typedef struct MCXBatchOpt {
    unsigned int blocksize;      /* lanes per workgroup: a multiple of 64 up to 1024 (each workgroup's
                                    LDS deposit box grows with its share of the CU's LDS); 0 picks the
                                    shape that keeps the most lanes resident, 256 lanes on ties */
    unsigned int noprobe;        /* 1: balance the per-XCD instance lists by photon counts instead of the
                                    costs a probe step measures */
    unsigned int usegraph;       /* 1: each step replays a captured HIP graph (prepare, transport, finalize) */
    unsigned int nqueue;         /* instance lists (list q runs on XCD q % 8); 0 picks 64, or fewer for
                                    small batches */
    /* filled in by the library */
    unsigned int nblock;         /* resident workgroups of the launch */
    unsigned int reserved;
    float stepms;                /* device time of the last step (prepare + transport + finalize) */
    float kernelms;              /* device time of the last transport kernel (steps run without a graph) */
    float tunems;                /* time spent choosing the schedule */
} MCXBatchOpt;

typedef struct MCXBatch MCXBatch;

/* Validate and stage n prepared Configs (as mcx_run_simulation takes them) on
   the device of the first; the Configs stay owned by the caller and must
   outlive the batch.  Restrictions: fast arithmetic (mathmode 0), no detected
   photons, no replay.  Returns 0, or an error id with mcx_last_error(). */
int mcx_batch_create(Config** cfgs, unsigned int n, MCXBatchOpt* opt, MCXBatch** out);
/* nstep steps; each simulates every instance afresh (the RNG streams continue) */
int mcx_batch_run(MCXBatch* batch, unsigned int nstep);
/* the last step's fluence into each Config's exportfield (allocated when NULL),
   its raw grid into exportfield64 when set, and energytot/energyesc/energyabs/
   normalizer/kernelms */
int mcx_batch_collect(MCXBatch* batch);
void mcx_batch_destroy(MCXBatch* batch);
/* create, one step, collect, destroy */
int mcx_run_batch(Config** cfgs, unsigned int n, MCXBatchOpt* opt);

#ifdef __cplusplus
}
#endif

#endif

/*
 * mcx_core_fast.hip - transport kernel instances using gfx950 hardware
 * transcendentals and reciprocal-based division (the analogue of the
 * reference's -DMCX_USE_NATIVE build).  Built without floating-point
 * contraction: the compiler fuses multiply-adds differently in each template
 * instance, and a replay instance must round the forward instance's path
 * arithmetic identically to retrace its photons.
 */
#define MCX_PARITY 0
#include "mcx_core_impl.h"

template <bool SAVEDET, bool REFLECT>
static hipError_t launch_variant(const MCXKParam* prm, int nblock, int nthreadblock, size_t shmem, hipStream_t stream) {
    if (prm->simple) {
        hipLaunchKernelGGL((mcxk::transport_kernel<SAVEDET, REFLECT, true, false>), dim3(nblock), dim3(nthreadblock),
                           shmem, stream, *prm);
    } else if (prm->full) {
        hipLaunchKernelGGL((mcxk::transport_kernel<SAVEDET, REFLECT, false, true>), dim3(nblock), dim3(nthreadblock),
                           shmem, stream, *prm);
    } else {
        hipLaunchKernelGGL((mcxk::transport_kernel<SAVEDET, REFLECT, false, false>), dim3(nblock), dim3(nthreadblock),
                           shmem, stream, *prm);
    }

    return hipGetLastError();
}

extern "C" hipError_t mcx_launch_fast(const MCXKParam* prm, int nblock, int nthreadblock, size_t shmem,
                                      hipStream_t stream, int savedet, int reflect) {
    if (savedet) {
        return reflect ? launch_variant<true, true>(prm, nblock, nthreadblock, shmem, stream)
               : launch_variant<true, false>(prm, nblock, nthreadblock, shmem, stream);
    }

    return reflect ? launch_variant<false, true>(prm, nblock, nthreadblock, shmem, stream)
           : launch_variant<false, false>(prm, nblock, nthreadblock, shmem, stream);
}

/* Resident workgroups per CU of a kernel from its own register and LDS use and
   gfx950's per-CU resources: 4 SIMDs, each holding at most 8 waves and 512
   VGPRs per lane allocated in granules of 8, and the device's LDS per CU in
   512-byte granules per workgroup.  A workgroup spreads its waves over the 4
   SIMDs.  (The runtime's occupancy query has returned half of this count on
   some MI355X hosts for the same code object, ruling out 1024-lane workgroups
   that launch and run there.) */
static int gfx950_occupancy(const void* kernel, int nthreadblock, size_t shmem) {
    hipFuncAttributes a;
    int dev = 0, ldscu = 0;

    if (nthreadblock <= 0 || hipFuncGetAttributes(&a, kernel) != hipSuccess || hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ldscu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess ||
            nthreadblock > a.maxThreadsPerBlock || ldscu <= 0) {
        return 0;
    }

    const int waves = (nthreadblock + 63) / 64;
    const int perSimd = (waves + 3) / 4;
    const int vgpr = ((a.numRegs > 0 ? a.numRegs : 1) + 7) & ~7;
    const int simdWaves = (512 / vgpr) < 8 ? (512 / vgpr) : 8;
    const int byRegs = simdWaves / perSimd;
    const size_t lds = (shmem + a.sharedSizeBytes + 511) & ~(size_t)511;
    const int byLds = lds ? (int)((size_t)ldscu / lds) : byRegs;
    return byRegs < byLds ? byRegs : byLds;
}

/* the transport kernel instance of a variant: path 1 the pencil-beam
   specialisation, 0 the general path, 2 the full general path */
template <bool SAVEDET, bool REFLECT>
static const void* kernel_of(int path) {
    return path == 1 ? (const void*)mcxk::transport_kernel<SAVEDET, REFLECT, true, false> :
           (path == 2 ? (const void*)mcxk::transport_kernel<SAVEDET, REFLECT, false, true> :
            (const void*)mcxk::transport_kernel<SAVEDET, REFLECT, false, false>);
}

static const void* transport_kernel_of(int savedet, int reflect, int path) {
    return savedet ? (reflect ? kernel_of<true, true>(path) : kernel_of<true, false>(path)) :
           (reflect ? kernel_of<false, true>(path) : kernel_of<false, false>(path));
}

extern "C" int mcx_fast_occupancy(int nthreadblock, size_t shmem, int savedet, int reflect, int simple) {
    return gfx950_occupancy(transport_kernel_of(savedet, reflect, simple), nthreadblock, shmem);
}

/* ---------------- batched launch ---------------- */
extern "C" hipError_t mcx_launch_batch(const MCXKParam* prm, const uint32_t* qlist, uint32_t nqueue, uint32_t* state,
                                       uint32_t* nseed, double* energy, unsigned long long* cost, int nblock,
                                       int nthreadblock, size_t shmem, int reflect, hipStream_t stream) {
    if (reflect) {
        hipLaunchKernelGGL(mcxk::batch_kernel<true>, dim3(nblock), dim3(nthreadblock), shmem, stream, prm, qlist,
                           nqueue, state, nseed, energy, cost);
    } else {
        hipLaunchKernelGGL(mcxk::batch_kernel<false>, dim3(nblock), dim3(nthreadblock), shmem, stream, prm, qlist,
                           nqueue, state, nseed, energy, cost);
    }

    return hipGetLastError();
}

extern "C" int mcx_batch_occupancy(int nthreadblock, size_t shmem, int reflect) {
    return reflect ? gfx950_occupancy((const void*)mcxk::batch_kernel<true>, nthreadblock, shmem) :
           gfx950_occupancy((const void*)mcxk::batch_kernel<false>, nthreadblock, shmem);
}

/* the kernel attributes the occupancy model reads (VGPRs, static LDS, lane
   bound), for tests */
extern "C" int mcx_fast_kernel_attr(int savedet, int reflect, int simple, int* vgpr, int* maxlanes) {
    hipFuncAttributes a;
    const void* k = transport_kernel_of(savedet, reflect, simple);

    if (hipFuncGetAttributes(&a, k) != hipSuccess) {
        return -1;
    }

    *vgpr = a.numRegs;
    *maxlanes = a.maxThreadsPerBlock;
    return 0;
}

/* start of a batched step: zero the grids and energy tallies and restore the
   schedule state (list heads and photon counters) from its initial copy; a
   kernel, not hipMemsetAsync, so the step can be captured in a HIP graph and
   replayed */
__global__ void batch_prep_kernel(double* __restrict__ field, size_t nfield, double* __restrict__ energy,
                                  uint32_t nenergy, uint32_t* __restrict__ state, const uint32_t* __restrict__ state0,
                                  uint32_t nstate) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t i0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x;

    for (size_t i = i0; i < nfield; i += stride) {
        field[i] = 0.0;
    }

    for (size_t i = i0; i < nenergy; i += stride) {
        energy[i] = 0.0;
    }

    for (size_t i = i0; i < nstate; i += stride) {
        state[i] = state0[i];
    }
}

extern "C" hipError_t mcx_launch_batch_prep(double* field, size_t nfield, double* energy, uint32_t nenergy,
        uint32_t* state, const uint32_t* state0, uint32_t nstate, hipStream_t stream) {
    size_t need = nfield > nenergy ? nfield : nenergy;
    need = need > nstate ? need : nstate;
    size_t nb = (need + 255) / 256;
    nb = nb < 4096 ? (nb ? nb : 1) : 4096;
    hipLaunchKernelGGL(batch_prep_kernel, dim3((unsigned int)nb), dim3(256), 0, stream, field, nfield, energy,
                       nenergy, state, state0, nstate);
    return hipGetLastError();
}

/* end of a batched step: each instance's float output from its grid, scaled by
   the normalisation of its output type with the step's launched energy
   (the single-run host's formulas, src/mcx_host.cpp:1382-1465); blockIdx.y
   selects the instance */
__global__ void batch_finalize_kernel(const MCXBatchNorm* __restrict__ norm, const double* __restrict__ energy) {
    const MCXBatchNorm& b = norm[blockIdx.y];
    const double etot = energy[2 * blockIdx.y + 1];
    float scale = 1.f;

    if (b.normalize && etot > 0.0) {
        if (b.outputtype == 0 || b.outputtype == 1) {
            scale = b.unitinmm / (float)(etot * (double)b.vvox * (double)b.tstep);

            if (b.outputtype == 1) {
                scale *= b.tstep;
            }
        } else if (b.outputtype == 2 || b.outputtype == 7) {
            scale = 1.f / (float)etot;
        }
    }

    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < b.len; i += (size_t)gridDim.x * blockDim.x) {
        b.out[i] = (float)b.field[i] * scale;
    }
}

extern "C" hipError_t mcx_launch_batch_finalize(const MCXBatchNorm* norm, const double* energy, uint32_t ninst,
        size_t maxlen, hipStream_t stream) {
    size_t nb = (maxlen + 255) / 256;
    nb = nb < 1024 ? (nb ? nb : 1) : 1024;
    hipLaunchKernelGGL(batch_finalize_kernel, dim3((unsigned int)nb, ninst), dim3(256), 0, stream, norm, energy);
    return hipGetLastError();
}

/* device self-check: math in fast (hardware) mode */
__global__ void fpmath_fast_kernel(int which, const float* __restrict__ x, float* __restrict__ y, size_t n) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;

    if (i >= n) {
        return;
    }

    float a = x[i], r = 0.f, s, c;

    switch (which) {
        case 0:
            r = mcxk::kexp(a);
            break;

        case 1:
            r = mcxk::klog(a);
            break;

        case 2:
            mcxk::ksincos(a, &s, &c);
            r = s;
            break;

        case 3:
            mcxk::ksincos(a, &s, &c);
            r = c;
            break;

        case 4:
            r = mcxk::kacos(a);
            break;

        case 5:
            r = mcxk::krsqrt(a);
            break;

        case 6:
            r = mcxk::kdiv(a, x[i + n]);
            break;

        default:
            r = mcxk::ksqrt(a);
            break;
    }

    y[i] = r;
}

extern "C" hipError_t mcx_launch_fpmath_fast(int which, const float* x, float* y, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(fpmath_fast_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, which, x, y, n);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
}

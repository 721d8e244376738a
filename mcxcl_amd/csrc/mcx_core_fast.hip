/*
 * mcx_core_fast.hip - transport kernel instances using gfx950 hardware
 * transcendentals and reciprocal-based division (the analogue of the
 * reference's -cl-mad-enable -DMCX_USE_NATIVE build).
 */
#define MCX_PARITY 0
#include "mcx_core_impl.h"

template <bool SAVEDET, bool REFLECT>
static hipError_t launch_variant(const MCXKParam* prm, int nblock, int nthreadblock, size_t shmem, hipStream_t stream) {
    if (prm->simple) {
        hipLaunchKernelGGL((mcxk::transport_kernel<SAVEDET, REFLECT, true>), dim3(nblock), dim3(nthreadblock), shmem,
                           stream, *prm);
    } else {
        hipLaunchKernelGGL((mcxk::transport_kernel<SAVEDET, REFLECT, false>), dim3(nblock), dim3(nthreadblock), shmem,
                           stream, *prm);
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

template <bool SAVEDET, bool REFLECT, bool SIMPLE>
static int occupancy_of(int nthreadblock, size_t shmem) {
    int nb = 0;

    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, mcxk::transport_kernel<SAVEDET, REFLECT, SIMPLE>,
            nthreadblock, shmem) != hipSuccess) {
        return 0;
    }

    return nb;
}

extern "C" int mcx_fast_occupancy(int nthreadblock, size_t shmem, int savedet, int reflect, int simple) {
    if (savedet) {
        if (reflect) {
            return simple ? occupancy_of<true, true, true>(nthreadblock, shmem)
                   : occupancy_of<true, true, false>(nthreadblock, shmem);
        }

        return simple ? occupancy_of<true, false, true>(nthreadblock, shmem)
               : occupancy_of<true, false, false>(nthreadblock, shmem);
    }

    if (reflect) {
        return simple ? occupancy_of<false, true, true>(nthreadblock, shmem)
               : occupancy_of<false, true, false>(nthreadblock, shmem);
    }

    return simple ? occupancy_of<false, false, true>(nthreadblock, shmem)
           : occupancy_of<false, false, false>(nthreadblock, shmem);
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

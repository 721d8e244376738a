/*
 * mcx_core_parity.hip - transport kernel instances with bit-reproducible
 * arithmetic (deterministic FP32 math, IEEE divide/sqrt).  Must be compiled
 * with -ffp-contract=off so no multiply-add is fused.
 */
#define MCX_PARITY 1
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

extern "C" hipError_t mcx_launch_parity(const MCXKParam* prm, int nblock, int nthreadblock, size_t shmem,
                                        hipStream_t stream, int savedet, int reflect) {
    if (savedet) {
        return reflect ? launch_variant<true, true>(prm, nblock, nthreadblock, shmem, stream)
               : launch_variant<true, false>(prm, nblock, nthreadblock, shmem, stream);
    }

    return reflect ? launch_variant<false, true>(prm, nblock, nthreadblock, shmem, stream)
           : launch_variant<false, false>(prm, nblock, nthreadblock, shmem, stream);
}

/* RNG stream dump (reference debug mode -D R, src/mcx_core.cl:2408-2414) */
__global__ void rngdump_kernel(const MCXKParam g) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;

    if (idx >= g.nthread) {
        return;
    }

    mcxk::Rng rng;
    rng.s0 = (uint64_t)g.nseed[idx * 4] << 32 | g.nseed[idx * 4 + 1];
    rng.s1 = (uint64_t)g.nseed[idx * 4 + 2] << 32 | g.nseed[idx * 4 + 3];

    for (uint32_t i = idx; i < g.fieldlen; i += g.nthread) {
        g.field[i] = (double)rng.next();
    }
}

extern "C" hipError_t mcx_launch_rngdump(const MCXKParam* prm, int nblock, int nthreadblock, hipStream_t stream) {
    hipLaunchKernelGGL(rngdump_kernel, dim3(nblock), dim3(nthreadblock), 0, stream, *prm);
    return hipGetLastError();
}

/* sum the accumulator replicas (fixed order) into replica 0, then write the
   normalised float output; grid-stride */
__global__ void finalize_kernel(double* __restrict__ in, float* __restrict__ out, size_t n, unsigned int nrep,
                                float scale) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double acc = in[i];

        for (unsigned int r = 1; r < nrep; r++) {
            acc += in[(size_t)r * n + i];
        }

        in[i] = acc;

        if (out) {   /* NULL: only fold the replicas */
            float v = (float)acc;
            out[i] = v * scale;
        }
    }
}

extern "C" hipError_t mcx_launch_finalize(double* in, float* out, size_t n, unsigned int nrep, float scale,
                                          hipStream_t stream) {
    int nblock = (int)((n + 255) / 256);
    nblock = nblock > 8192 ? 8192 : (nblock < 1 ? 1 : nblock);
    hipLaunchKernelGGL(finalize_kernel, dim3(nblock), dim3(256), 0, stream, in, out, n, nrep, scale);
    return hipGetLastError();
}

/* ---- adjoint (phi_src . phi_det) outputs (src/mcx_core.cl:3313-3512) ----
   The normalised float field holds one slot of maxgate gates per source
   (user sources first, then one per detector) at field[vox + (t + slot *
   maxgate) * nvox].  Per voxel and per (source s, detector d) pair:
     kind 0: sum_t phi_s * sum_t phi_d
     kind 1: grad(sum_t phi_s) . grad(sum_t phi_d), second-order differences
             in voxel units (one-sided at the faces, first order on 2-voxel axes)
   times scale, then times opscale[label] when opscale is given.  Built
   without contraction: the gate sums run in gate order and every product and
   sum is rounded as written, so the CPU restatement matches bit for bit. */
__device__ __forceinline__ float cw_sum(const float* __restrict__ f, uint32_t vox, uint32_t slot, uint32_t maxgate,
                                        uint32_t nvox) {
    float acc = 0.f;

    for (uint32_t t = 0; t < maxgate; t++) {
        acc += f[vox + (size_t)(t + slot * maxgate) * nvox];
    }

    return acc;
}

__device__ __forceinline__ float fd_grad(const float* __restrict__ f, uint32_t vox, uint32_t slot, uint32_t maxgate,
        uint32_t nvox, uint32_t i, uint32_t n, uint32_t stride) {
    if (n <= 1) {
        return 0.f;
    }

    const float f0 = cw_sum(f, vox, slot, maxgate, nvox);

    if (i == 0) {
        const float p1 = cw_sum(f, vox + stride, slot, maxgate, nvox);

        if (n == 2) {
            return p1 - f0;
        }

        const float p2 = cw_sum(f, vox + 2 * stride, slot, maxgate, nvox);
        return (-3.f * f0 + 4.f * p1 - p2) * 0.5f;
    }

    if (i == n - 1) {
        const float m1 = cw_sum(f, vox - stride, slot, maxgate, nvox);

        if (n == 2) {
            return f0 - m1;
        }

        const float m2 = cw_sum(f, vox - 2 * stride, slot, maxgate, nvox);
        return (m2 - 4.f * m1 + 3.f * f0) * 0.5f;
    }

    return (cw_sum(f, vox + stride, slot, maxgate, nvox) - cw_sum(f, vox - stride, slot, maxgate, nvox)) * 0.5f;
}

/* the adjoint outputs of one voxel for every (source, detector) pair: kind 0
   sum_t phi_s * sum_t phi_d (src/mcx_core.cl:3466-3512), kind 1 grad phi_s .
   grad phi_d (:3390-3464); complex for RF fields (im != NULL): re = re.re -
   im.im, im = re.im + im.re, the imaginary part to out_im.  Both times scale
   and the label's factor. */
__global__ void adjoint_kernel(const float* __restrict__ field, const float* __restrict__ field_im,
                               float* __restrict__ out, float* __restrict__ out_im, uint32_t nvox, uint32_t nx,
                               uint32_t ny, uint32_t maxgate, uint32_t ns, uint32_t nd, int kind, float scale,
                               const float* __restrict__ opscale, const uint32_t* __restrict__ media) {
    const uint32_t vox = blockIdx.x * blockDim.x + threadIdx.x;

    if (vox >= nvox) {
        return;
    }

    const uint32_t nxy = nx * ny, nz = nvox / nxy;
    const uint32_t ix = vox % nx, iy = (vox / nx) % ny, iz = vox / nxy;
    const float op = opscale ? opscale[media[vox] & 0xFFu] : 1.f;

    for (uint32_t s = 0; s < ns; s++) {
        float sx = 0.f, sy = 0.f, sz = 0.f, sc = 0.f;
        float sxi = 0.f, syi = 0.f, szi = 0.f, sci = 0.f;

        if (kind == 0) {
            sc = cw_sum(field, vox, s, maxgate, nvox);

            if (field_im) {
                sci = cw_sum(field_im, vox, s, maxgate, nvox);
            }
        } else {
            sx = fd_grad(field, vox, s, maxgate, nvox, ix, nx, 1);
            sy = fd_grad(field, vox, s, maxgate, nvox, iy, ny, nx);
            sz = fd_grad(field, vox, s, maxgate, nvox, iz, nz, nxy);

            if (field_im) {
                sxi = fd_grad(field_im, vox, s, maxgate, nvox, ix, nx, 1);
                syi = fd_grad(field_im, vox, s, maxgate, nvox, iy, ny, nx);
                szi = fd_grad(field_im, vox, s, maxgate, nvox, iz, nz, nxy);
            }
        }

        for (uint32_t d = 0; d < nd; d++) {
            float v, vi = 0.f;

            if (kind == 0) {
                const float dc = cw_sum(field, vox, ns + d, maxgate, nvox);
                v = sc * dc;

                if (field_im) {
                    const float dci = cw_sum(field_im, vox, ns + d, maxgate, nvox);
                    v = v - sci * dci;
                    vi = sc * dci + sci * dc;
                }
            } else {
                const float dx = fd_grad(field, vox, ns + d, maxgate, nvox, ix, nx, 1);
                const float dy = fd_grad(field, vox, ns + d, maxgate, nvox, iy, ny, nx);
                const float dz = fd_grad(field, vox, ns + d, maxgate, nvox, iz, nz, nxy);
                v = sx * dx + sy * dy + sz * dz;

                if (field_im) {
                    const float dxi = fd_grad(field_im, vox, ns + d, maxgate, nvox, ix, nx, 1);
                    const float dyi = fd_grad(field_im, vox, ns + d, maxgate, nvox, iy, ny, nx);
                    const float dzi = fd_grad(field_im, vox, ns + d, maxgate, nvox, iz, nz, nxy);
                    v = v - (sxi * dxi + syi * dyi + szi * dzi);
                    vi = sx * dxi + sy * dyi + sz * dzi + sxi * dx + syi * dy + szi * dz;
                }
            }

            v = v * scale;
            vi = vi * scale;

            if (opscale) {
                v = v * op;
                vi = vi * op;
            }

            out[vox + (size_t)(s * nd + d) * nvox] = v;

            if (out_im) {
                out_im[vox + (size_t)(s * nd + d) * nvox] = vi;
            }
        }
    }
}

extern "C" hipError_t mcx_launch_adjoint(const float* field, const float* field_im, float* out, float* out_im,
        uint32_t nvox, uint32_t nx, uint32_t ny, uint32_t maxgate, uint32_t ns, uint32_t nd, int kind, float scale,
        const float* opscale, const uint32_t* media, hipStream_t stream) {
    hipLaunchKernelGGL(adjoint_kernel, dim3((nvox + 255) / 256), dim3(256), 0, stream, field, field_im, out, out_im,
                       nvox, nx, ny, maxgate, ns, nd, kind, scale, opscale, media);
    return hipGetLastError();
}

/* ---- device self-checks: math in parity mode, raw RNG stream ---- */
__global__ void fpmath_kernel(int which, const float* __restrict__ x, float* __restrict__ y, size_t n) {
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

extern "C" hipError_t mcx_launch_fpmath_parity(int which, const float* x, float* y, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(fpmath_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, which, x, y, n);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
}

__global__ void rngstream_kernel(const uint32_t* seed4, unsigned int n, float* out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) {
        return;
    }

    mcxk::Rng rng;
    rng.s0 = (uint64_t)seed4[0] << 32 | seed4[1];
    rng.s1 = (uint64_t)seed4[2] << 32 | seed4[3];

    for (unsigned int i = 0; i < n; i++) {
        out[i] = rng.next();
    }
}

/* throughput ceiling of float64 atomic adds at scattered addresses: every lane
   adds 1.0 to k hashed cells of an n-cell grid (the fluence grid's size) */
__global__ void atomic_probe_kernel(double* __restrict__ grid, uint32_t n, uint32_t k) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;

    for (uint32_t i = 0; i < k; i++) {
        uint32_t h = t * 2654435761u + i * 40503u;
        h ^= h >> 16;
        h *= 0x7FEB352Du;
        h ^= h >> 15;
        h *= 0x846CA68Bu;
        h ^= h >> 16;
        atomicAdd(grid + (h % n), 1.0);
    }
}

extern "C" hipError_t mcx_launch_atomic_probe(double* grid, uint32_t n, uint32_t nblock, uint32_t k,
        hipStream_t s) {
    hipLaunchKernelGGL(atomic_probe_kernel, dim3(nblock), dim3(256), 0, s, grid, n, k);
    return hipGetLastError();
}

extern "C" hipError_t mcx_launch_rngstream(const uint32_t* seed4, unsigned int n, float* out, hipStream_t s) {
    hipLaunchKernelGGL(rngstream_kernel, dim3(1), dim3(64), 0, s, seed4, n, out);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
}

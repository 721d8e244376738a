/*
 * mcx_core_impl.h - gfx950 photon-transport kernel body.
 *
 * Included by mcx_core_parity.hip (MCX_PARITY=1, built with
 * -ffp-contract=off: deterministic FP32 math from mcx_fpmath.h, IEEE divide
 * and sqrt, bit-identical to the CPU oracle) and by mcx_core_fast.hip
 * (MCX_PARITY=0: gfx950 hardware transcendentals v_exp/v_log/v_sin/v_cos/
 * v_rsq/v_rcp, the analogue of the reference's native_* build).
 *
 * Semantics: one persistent lane per reference thread, running its
 * threadphoton(+1) photons back to back on a per-lane xorshift128+ stream
 * (reference mcx_main_loop, src/mcx_core.cl:2307-3307, and launchnewphoton,
 * :1466-2275).  The RNG draws, their order and every single-precision
 * operation on the photon state match the reference's OpenCL branches.
 *
 * MI355X-specific structure:
 *  - one relaunch site: the reference inlines launchnewphoton at six sites;
 *    here every termination sets a flag and the loop head relaunches, so the
 *    large launch body is emitted once and lanes reconverge there;
 *  - the medium table and the phase-function inverse CDF are staged in LDS;
 *    the per-lane partial-path/energy scratch lives in LDS (wave64 stride);
 *  - fluence is accumulated with native global_atomic_add_f64 into a double
 *    grid (no MAX_ACCUM shadow half is needed), converted and normalised by a
 *    separate streaming kernel;
 *  - detected photons are compacted with one atomic per record.
 */
#include <float.h>

#include "mcx_core.h"
#include "mcx_fpmath.h"

#ifndef MCX_PARITY
#error "MCX_PARITY must be defined"
#endif

#define K_EPS          FLT_EPSILON
#define K_TWO_PI       6.28318530717959f
#define K_R_PI         0.318309886183791f
#define K_NANGLES      181u
#define K_ONE_PI       3.1415926535897932f
#define K_JUST_BELOW_ONE 0.9998f
#define K_ROULETTE     10.f
#define K_DET_MASK     0x80000000u
#define K_MED_MASK     0x7FFFFFFFu
#define K_OUT_MIN      0xFFFFFFFFu
#define K_OUT_MAX      0x7FFFFFFFu
#define K_SEED_FROM_FILE (-999)

enum {kReflect = 1, kAbsorb = 2, kMirror = 3, kCyclic = 4};
/* issavedet value of fill-max mode: stop once maxdetphoton photons are detected */
#define K_FILL_MAXDET 3u
enum {kFlux = 0, kFluence = 1, kEnergy = 2, kJacobian = 3, kWP = 4, kDCS = 5, kRF = 6, kL = 7, kRFmus = 8,
      kWLTOF = 9, kWPTOF = 10
     };
enum {kPencil = 0, kIsotropic = 1, kCone = 2, kGaussian = 3, kPlanar = 4, kPattern = 5, kFourier = 6, kArcsine = 7,
      kDisk = 8, kFourierX = 9, kFourierX2D = 10, kZGaussian = 11, kLine = 12, kSlit = 13, kPencilArray = 14,
      kPattern3D = 15, kHyperboloid = 16, kRing = 17
     };

/* Each build (parity, fast) gets its own namespace: the kernels are templates
   instantiated with the same arguments in both translation units, and under
   one name their host launch stubs would be merged at link time, so a launch
   from one build could run the other build's code object. */
#if MCX_PARITY
#define MCXK_NS mcxk_parity
#else
#define MCXK_NS mcxk_fast
#endif

namespace MCXK_NS {

#if MCX_PARITY
__device__ __forceinline__ float kexp(float x) {
    return pm_exp(x);
}
__device__ __forceinline__ float klog(float x) {
    return pm_log(x);
}
__device__ __forceinline__ void ksincos(float x, float* s, float* c) {
    pm_sincos(x, s, c);
}
__device__ __forceinline__ float ksin(float x) {
    return pm_sin(x);
}
__device__ __forceinline__ float kcos(float x) {
    return pm_cos(x);
}
__device__ __forceinline__ float kacos(float x) {
    return pm_acos(x);
}
__device__ __forceinline__ float ksqrt(float x) {
    return __builtin_sqrtf(x);
}
__device__ __forceinline__ float krsqrt(float x) {
    return 1.f / __builtin_sqrtf(x);
}
__device__ __forceinline__ float kdiv(float a, float b) {
    return a / b;
}
#else
__device__ __forceinline__ float kexp(float x) {
    return __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
}
__device__ __forceinline__ float klog(float x) {
    return __builtin_amdgcn_logf(x) * 0.693147180559945309f;
}
__device__ __forceinline__ void ksincos(float x, float* s, float* c) {
    float r = x * 0.159154943091895336f;   /* v_sin/v_cos take revolutions */
    *s = __builtin_amdgcn_sinf(r);
    *c = __builtin_amdgcn_cosf(r);
}
__device__ __forceinline__ float ksin(float x) {
    return __builtin_amdgcn_sinf(x * 0.159154943091895336f);
}
__device__ __forceinline__ float kcos(float x) {
    return __builtin_amdgcn_cosf(x * 0.159154943091895336f);
}
__device__ __forceinline__ float ksqrt(float x) {
    return __builtin_amdgcn_sqrtf(x);
}
__device__ __forceinline__ float krsqrt(float x) {
    return __builtin_amdgcn_rsqf(x);
}
__device__ __forceinline__ float kdiv(float a, float b) {
    return a * __builtin_amdgcn_rcpf(b);
}
__device__ __forceinline__ float kacos(float x) {
    /* same Cephes kernel as parity mode, with the hardware square root */
    const float PI = 3.14159265358979f, PIO2 = 1.57079632679490f;

    if (x > 0.5f) {
        return 2.f * pm_asin_kernel(__builtin_amdgcn_sqrtf(0.5f * (1.f - x)));
    }

    if (x < -0.5f) {
        return PI - 2.f * pm_asin_kernel(__builtin_amdgcn_sqrtf(0.5f * (1.f + x)));
    }

    return PIO2 - pm_asin_kernel(x);
}
#endif

/* sin(acos(x)) for |x| <= 1.  Parity mode evaluates it as the reference writes
   it; fast mode uses the identity sqrt(1 - x^2). */
__device__ __forceinline__ float ksin_acos(float x) {
#if MCX_PARITY
    return ksin(kacos(x));
#else
    return __builtin_amdgcn_sqrtf(fmaxf(0.f, 1.f - x * x));
#endif
}

/* LDS deposit box cells: float64 in the parity build, float32 in the fast
   build (twice the cells for the same LDS; the reference accumulates its whole
   grid in float32).  The host sizes the box with the same element size. */
#if MCX_PARITY
typedef double BoxCell;
#else
typedef float BoxCell;
#endif

/* ---------------- RNG: xorshift128+ (src/mcx_core.cl:684-698) ---------------- */
struct Rng {
    uint64_t s0, s1;
    __device__ __forceinline__ float next() {
        uint64_t x = s0;
        const uint64_t y = s1;
        s0 = y;
        x ^= x << 23;
        s1 = x ^ y ^ (x >> 18) ^ (y >> 5);
        uint32_t lo = (uint32_t)(s1 + y);
        return __uint_as_float(0x3F800000u | (lo >> 9)) - 1.0f;
    }
    __device__ __forceinline__ float scatlen() {
        return -klog(next() + K_EPS);
    }
};

__device__ __forceinline__ int to_short(int v) {
    return (int)(short)v;
}

__device__ __forceinline__ bool inside(const MCXKParam& g, int x, int y, int z) {
    return (uint32_t)(x & 0xFFFF) < g.dimx && (uint32_t)(y & 0xFFFF) < g.dimy && (uint32_t)(z & 0xFFFF) < g.dimz;
}

__device__ __forceinline__ float nextafter_ref(float a, int dir) {
    uint32_t u = __float_as_uint(a + 1000.f);
    u += (uint32_t)(dir ^ (int)(u & 0x80000000u));
    return __uint_as_float(u) - 1000.f;
}

/* distance to the next voxel face and the face id (src/mcx_core.cl:975-995) */
__device__ __forceinline__ float hitgrid(const float4& p, const float4& v, int ix, int iy, int iz, int* face) {
    float hx = fabsf((float)ix + (float)(v.x > 0.f) - p.x);
    float hy = fabsf((float)iy + (float)(v.y > 0.f) - p.y);
    float hz = fabsf((float)iz + (float)(v.z > 0.f) - p.z);
    hx = fabsf(kdiv(hx + K_EPS, v.x));
    hy = fabsf(kdiv(hy + K_EPS, v.y));
    hz = fabsf(kdiv(hz + K_EPS, v.z));
    float dist = fminf(fminf(hx, hy), hz);
    *face = (dist == hx ? 0 : (dist == hy ? 1 : 2));
    return dist;
}

/* src/mcx_core.cl:1025-1042, including the three sequential renormalisations */
__device__ __forceinline__ void rotatevector(float4& v, float stheta, float ctheta, float sphi, float cphi) {
    if (v.z > -1.f + K_EPS && v.z < 1.f - K_EPS) {
        float tmp0 = 1.f - v.z * v.z;
        float tmp1 = stheta * krsqrt(tmp0);
        float nx = tmp1 * (v.x * v.z * cphi - v.y * sphi) + v.x * ctheta;
        float ny = tmp1 * (v.y * v.z * cphi + v.x * sphi) + v.y * ctheta;
        float nz = -tmp1 * tmp0 * cphi + v.z * ctheta;
        v.x = nx;
        v.y = ny;
        v.z = nz;
    } else {
        float nx = stheta * cphi, ny = stheta * sphi, nz = (v.z > 0.f) ? ctheta : -ctheta;
        v.x = nx;
        v.y = ny;
        v.z = nz;
    }

#if MCX_PARITY
    /* three sequential renormalisations, as the reference writes them */
    v.x *= krsqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    v.y *= krsqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    v.z *= krsqrt(v.x * v.x + v.y * v.y + v.z * v.z);
#else
    float rn = krsqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    v.x *= rn;
    v.y *= rn;
    v.z *= rn;
#endif
}

__device__ __forceinline__ void rotatevector2d(float4& v, float stheta, float ctheta, int is2d) {
    float nx = v.x, ny = v.y, nz = v.z;

    if (is2d == 1) {
        nx = 0.f;
        ny = v.y * ctheta - v.z * stheta;
        nz = v.y * stheta + v.z * ctheta;
    } else if (is2d == 2) {
        nx = v.x * ctheta - v.z * stheta;
        ny = 0.f;
        nz = v.x * stheta + v.z * ctheta;
    } else if (is2d == 3) {
        nx = v.x * ctheta - v.y * stheta;
        ny = v.x * stheta + v.y * ctheta;
        nz = 0.f;
    }

    v.x = nx;
    v.y = ny;
    v.z = nz;
    float tmp0 = krsqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    v.x *= tmp0;
    v.y *= tmp0;
    v.z *= tmp0;
}

__device__ __forceinline__ void transmit(float4& v, float n1, float n2, int flipdir) {
    float tmp0 = kdiv(n1, n2);
    v.x *= tmp0;
    v.y *= tmp0;
    v.z *= tmp0;

    if (flipdir == 0) {
        v.x = ksqrt(1.f - v.y * v.y - v.z * v.z) * (float)((v.x > 0.f) - (v.x < 0.f));
    } else if (flipdir == 1) {
        v.y = ksqrt(1.f - v.x * v.x - v.z * v.z) * (float)((v.y > 0.f) - (v.y < 0.f));
    } else {
        v.z = ksqrt(1.f - v.x * v.x - v.y * v.y) * (float)((v.z > 0.f) - (v.z < 0.f));
    }
}

__device__ __forceinline__ float reflectcoeff(const float4& v, float n1, float n2, int flipdir) {
    float Icos = fabsf((flipdir == 0) ? v.x : (flipdir == 1 ? v.y : v.z));
    float tmp0 = n1 * n1;
    float tmp1 = n2 * n2;
    float tmp2 = 1.f - kdiv(tmp0, tmp1) * (1.f - Icos * Icos);

    if (tmp2 > 0.f) {
        float Re = tmp0 * Icos * Icos + tmp1 * tmp2;
        tmp2 = ksqrt(tmp2);
        float Im = 2.f * n1 * n2 * Icos * tmp2;
        float Rtotal = kdiv(Re - Im, Re + Im);
        Re = tmp1 * Icos * Icos + tmp0 * tmp2 * tmp2;
        return (Rtotal + kdiv(Re - Im, Re + Im)) * 0.5f;
    }

    return 1.f;
}

/* per-lane photon state */
struct Photon {
    float4 p, v, f, prop;
    int fx, fy, fz, fw;          /* voxel index (short semantics) and hit face */
    uint32_t idx1d, mediaid;
    float w0, Lmove;
    /* split voxel (svmc, src/mcx_core.cl:594-598): the interface plane n.x = pd of
       the current voxel, its normal pointing away from the photon's side, as
       {nx, ny, nz, pd} in the lane's scratch at svoff (sn points there in svmc
       runs: read on each step, but only by those runs, so it holds no registers
       in the others), and lower label | upper label << 8 | split << 16 | on the
       upper side << 17 */
    float* sn;
    uint32_t sv;
};

struct Shared {
    float4* prop;                /* medium table */
    float* invcdf;
    BoxCell* box;                /* privatised deposit box (box_cells) */
    float* pp;                   /* this lane's scratch: partial data, energies, w0, srcid */
    size_t repoff;               /* offset of this lane's global accumulator replica */
    uint32_t* pool;              /* dynamic schedule: {next photon, end} of this workgroup (LDS)
                                    or of this instance (global memory, batched launch) */
};

/* Per-lane write-combining buffer for grid deposits: the two most recent grid
   addresses and their pending sums.  A photon bouncing between two voxels
   deposits A, B, A, B, ...; those merge here and reach the grid as one atomic
   per eviction (13% of grid deposits on the 60^3 sphere workload).  Sums are
   kept in double, so merging two float deposits is exact. */
struct DepCache {
    uint32_t a0, a1;             /* addresses (within the replica), 0xFFFFFFFF = empty */
    double v0, v1;
};

__device__ __forceinline__ void depcache_add(const MCXKParam& g, const Shared& sh, DepCache& c, uint32_t addr,
        float val) {
    if (addr == c.a0) {
        c.v0 += (double)val;
    } else if (addr == c.a1) {
        c.a1 = c.a0;
        c.a0 = addr;
        double t = c.v1 + (double)val;
        c.v1 = c.v0;
        c.v0 = t;
    } else {
        if (c.a1 != 0xFFFFFFFFu) {
            atomicAdd(g.field + sh.repoff + c.a1, c.v1);
        }

        c.a1 = c.a0;
        c.v1 = c.v0;
        c.a0 = addr;
        c.v0 = (double)val;
    }
}

__device__ __forceinline__ void depcache_flush(const MCXKParam& g, const Shared& sh, DepCache& c) {
    if (c.a0 != 0xFFFFFFFFu) {
        atomicAdd(g.field + sh.repoff + c.a0, c.v0);
    }

    if (c.a1 != 0xFFFFFFFFu) {
        atomicAdd(g.field + sh.repoff + c.a1, c.v1);
    }

    c.a0 = c.a1 = 0xFFFFFFFFu;
}

/* voxel-exit deposit: LDS box when the voxel and gate fall inside it, else the
   lane's write-combining buffer in front of the global grid.

   A float32 box cell (fast build) keeps its magnitude below K_BOX_MOVE: the
   lane whose add finds the cell above it moves K_BOX_MOVE (exactly
   representable) from the cell to the float64 grid, as the reference moves
   MAX_ACCUM from its float32 grid to the shadow half (src/mcx_core.cl:2882-2886).
   Deposits therefore keep the rounding of a float32 sum below 2^11 however
   many photons a workgroup runs, and the moved amount is exact, so nothing is
   lost or counted twice when several lanes move at once. */
#define K_BOX_MOVE 1024.f
__device__ __forceinline__ void deposit_voxel(const MCXKParam& g, const Shared& sh, DepCache& c, int ox, int oy,
        int oz, uint32_t idx1d, int tshift, float val) {
    uint32_t dx = (uint32_t)(ox - (int)g.box_x0), dy = (uint32_t)(oy - (int)g.box_y0);
    uint32_t dz = (uint32_t)(oz - (int)g.box_z0);

    if (dx < g.box_nx && dy < g.box_ny && dz < g.box_nz && (uint32_t)tshift < g.box_ng) {
        BoxCell* cell = sh.box + (((uint32_t)tshift * g.box_nz + dz) * g.box_ny + dy) * g.box_nx + dx;
#if MCX_PARITY
        atomicAdd(cell, (BoxCell)val);
#else
        const float old = atomicAdd(cell, val);

        if (fabsf(old) > K_BOX_MOVE) {
            const float mv = old > 0.f ? K_BOX_MOVE : -K_BOX_MOVE;
            atomicAdd(cell, -mv);
            atomicAdd(g.field + sh.repoff + idx1d + (uint32_t)tshift * g.nvox, (double)mv);
        }
#endif
    } else {
        depcache_add(g, sh, c, idx1d + (uint32_t)tshift * g.nvox, val);
    }
}

/* binary16 bit pattern (sign bit clear) to float, exactly, subnormals included */
__device__ __forceinline__ float half_bits_to_float(uint32_t h) {
    const uint32_t e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;

    if (e == 0) {
        return (float)m * 5.9604644775390625e-8f;   /* m x 2^-24 */
    }

    return __uint_as_float(e == 31 ? (0x7F800000u | (m << 13)) : (((e + 112u) << 23) | (m << 13)));
}

/* Properties of a voxel whose volume word packs them instead of naming a label
   (src/mcx_core.cl:1079-1193).  t is the label table: t[0] the background,
   t[1] and t[2] the two end members of the mixing formats.  The components a
   format leaves unset keep t[1]'s, the value the reference's launch writes
   (src/mcx_core.cl:2213) and its updates never overwrite; n is t[1]'s in a
   non-void voxel.  Bit 31 (the detector mask) is not part of the word. */
enum {kMediaLabelHalf = 99, kMediaAsF2H = 100, kMediaMuaFloat = 101, kMediaAsHalf = 102, kMediaAsgnByte = 103,
      kMediaAsShort = 104};

__device__ __forceinline__ float4 media_word_prop(uint32_t format, const float4* t, uint32_t mediaid) {
    const uint32_t w = mediaid & K_MED_MASK;
    float4 p = t[1];
    const float nw = t[w != 0].w;

    if (format == kMediaMuaFloat) {
        p.x = __uint_as_float(w);                                  /* |mua| */
        p.w = nw;
    } else if (format == kMediaAsF2H || format == kMediaAsHalf) {
        p.x = half_bits_to_float(w & 0x7FFFu);                     /* |mua|, |mus| as two halves */
        p.y = half_bits_to_float(w >> 16);
        p.w = nw;
    } else if (format == kMediaLabelHalf) {
        p = t[w & 0x3FFFu];                                        /* label, then one component replaced */
        const float r = half_bits_to_float(w >> 16);
        const uint32_t k = (w >> 14) & 3u;
        p.x = (k == 0) ? r : p.x;
        p.y = (k == 1) ? r : p.y;
        p.z = (k == 2) ? r : p.z;
        p.w = (k == 3) ? r : p.w;
    } else if (format == kMediaAsgnByte) {
        p.x = (float)(w & 0xFFu) * (1.f / 255.f) * (t[2].x - t[1].x) + t[1].x;
        p.y = (float)((w >> 8) & 0xFFu) * (1.f / 255.f) * (t[2].y - t[1].y) + t[1].y;
        p.z = (float)((w >> 16) & 0xFFu) * (1.f / 255.f) * (t[2].z - t[1].z) + t[1].z;
        p.w = (float)(w >> 24) * (1.f / 127.f) * (t[2].w - t[1].w) + t[1].w;
    } else if (format == kMediaAsShort) {
        p.x = (float)(w & 0xFFFFu) * (1.f / 65535.f) * (t[2].x - t[1].x) + t[1].x;
        p.y = (float)(w >> 16) * (1.f / 65535.f) * (t[2].y - t[1].y) + t[1].y;
        p.w = nw;
    }

    return p;
}

/* ---- split-voxel media (svmc, src/mcx_core.cl:1229-1348) ---- */
#define K_SV_LOWER(sv)   ((sv) & 0xFFu)
#define K_SV_UPPER(sv)   (((sv) >> 8) & 0xFFu)
#define K_SV_ISSPLIT(sv) (((sv) >> 16) & 1u)
#define K_SV_ISUPPER(sv) (((sv) >> 17) & 1u)
#define K_SV_CURLABEL(sv) (K_SV_ISUPPER(sv) ? K_SV_UPPER(sv) : K_SV_LOWER(sv))
enum {kMediaSvmc = 97};

/* the properties of a split voxel at position p (its integer corner fx, fy, fz):
   the second volume word holds the plane normal (three bytes, 2c/255 - 1 each)
   and the reference point's z, the first word the reference point's y and x and
   the upper and lower labels; a voxel whose upper label is 0 is not split.  The
   side is the sign of p.n - pd; on the upper side the normal is turned to face
   the lower one.  Outside the grid: the background. */
__device__ __forceinline__ float4 svmc_prop(const MCXKParam& g, const float4* t, uint32_t mediaid, uint32_t idx1d,
        const float4& p, int fx, int fy, int fz, Photon& s) {
    if (idx1d == K_OUT_MIN || idx1d == K_OUT_MAX) {
        return t[0];
    }

    const uint32_t lo = g.media[idx1d + g.nvox];           /* bytes c0..c3: nz, ny, nx, pz */
    const uint32_t hi = mediaid & K_MED_MASK;                /* bytes c4..c7: py, px, upper, lower */
    const uint32_t c6 = (hi >> 16) & 0xFFu, c7 = hi >> 24;
    uint32_t sv = c7 | (c6 << 8);
    float4 prop;

    if (c6) {
        const float rx = (float)((hi >> 8) & 0xFFu) * (1.f / 255.f) + (float)fx;
        const float ry = (float)(hi & 0xFFu) * (1.f / 255.f) + (float)fy;
        const float rz = (float)(lo >> 24) * (1.f / 255.f) + (float)fz;
        float nx = (float)((lo >> 16) & 0xFFu) * (2.f / 255.f) - 1.f;
        float ny = (float)((lo >> 8) & 0xFFu) * (2.f / 255.f) - 1.f;
        float nz = (float)(lo & 0xFFu) * (2.f / 255.f) - 1.f;
        const float rn = krsqrt(nx * nx + ny * ny + nz * nz);
        nx = nx * rn;
        ny = ny * rn;
        nz = nz * rn;
        float pd = rx * nx + ry * ny + rz * nz;

        if (p.x * nx + p.y * ny + p.z * nz > pd) {
            prop = t[K_SV_UPPER(sv)];
            sv |= 1u << 17;
            nx = -nx;
            ny = -ny;
            nz = -nz;
            pd = -pd;
        } else {
            prop = t[K_SV_LOWER(sv)];
        }

        sv |= 1u << 16;
        s.sn[0] = nx;
        s.sn[1] = ny;
        s.sn[2] = nz;
        s.sn[3] = pd;
    } else {
        prop = t[c7];
    }

    s.sv = sv;
    return prop;
}

/* src/mcx_core.cl:1280-1299: where the step of length len crosses the voxel's
   plane towards the other side, shorten it to the plane and recompute the
   scattering length it uses; returns whether it does */
__device__ __forceinline__ bool svmc_ray_plane(const Photon& s, const float4& v, float gfac, float* len, float* slen) {
    const float vdotn = v.x * s.sn[0] + v.y * s.sn[1] + v.z * s.sn[2];

    if (vdotn <= 0.f) {
        return false;
    }

    const float d0 = s.p.x * s.sn[0] + s.p.y * s.sn[1] + s.p.z * s.sn[2] - s.sn[3];
    const float d1 = d0 + (*len) * vdotn;

    if (d0 * d1 > 0.f) {
        return false;
    }

    const float len0 = kdiv((*len) * d0, d0 - d1);
    *len = (len0 > 0.f) ? len0 : *len;
    *slen = (*len) * s.prop.y * gfac;
    return true;
}

/* src/mcx_core.cl:1300-1345: Fresnel reflection or refraction at the plane
   inside a split voxel, n1 the side the photon leaves; returns 1 when it
   refracts into label 0 (the photon leaves the medium) */
__device__ __forceinline__ int svmc_reflect(float n1, float4& v, Photon& s, const float4* t, Rng& rng) {
    float cx = v.x, cy = v.y, cz = v.z;
    const float icos = fabsf(cx * s.sn[0] + cy * s.sn[1] + cz * s.sn[2]);
    const float n2 = K_SV_ISUPPER(s.sv) ? t[K_SV_UPPER(s.sv)].w : t[K_SV_LOWER(s.sv)].w;
    const float tmp0 = n1 * n1, tmp1 = n2 * n2;
    float tmp2 = 1.f - kdiv(tmp0, tmp1) * (1.f - icos * icos);

    if (tmp2 > 0.f) {
        float re = tmp0 * icos * icos + tmp1 * tmp2;
        tmp2 = ksqrt(tmp2);
        const float im = 2.f * n1 * n2 * icos * tmp2;
        float rtotal = kdiv(re - im, re + im);
        re = tmp1 * icos * icos + tmp0 * tmp2 * tmp2;
        rtotal = (rtotal + kdiv(re - im, re + im)) * 0.5f;

        if (rng.next() <= rtotal) {
            const float k = -2.f * icos;
            cx += k * s.sn[0];
            cy += k * s.sn[1];
            cz += k * s.sn[2];
            s.sv ^= 1u << 17;
        } else {
            const float k = -icos, r = kdiv(n1, n2);
            cx += k * s.sn[0];
            cy += k * s.sn[1];
            cz += k * s.sn[2];
            cx = tmp2 * s.sn[0] + r * cx;
            cy = tmp2 * s.sn[1] + r * cy;
            cz = tmp2 * s.sn[2] + r * cz;
            s.sn[0] = -s.sn[0];
            s.sn[1] = -s.sn[1];
            s.sn[2] = -s.sn[2];
            s.sn[3] = -s.sn[3];

            if (K_SV_CURLABEL(s.sv) == 0) {
                return 1;
            }

            s.prop = t[K_SV_CURLABEL(s.sv)];
        }
    } else {
        const float k = -2.f * icos;
        cx += k * s.sn[0];
        cy += k * s.sn[1];
        cz += k * s.sn[2];
        s.sv ^= 1u << 17;
    }

    const float rn = krsqrt(cx * cx + cy * cy + cz * cz);
    v.x = cx * rn;
    v.y = cy * rn;
    v.z = cz * rn;
    return 0;
}

/* the properties of the photon's voxel: label table, per-voxel mua/mus
   replacing the label's, or (full general path) a packed media word */
template <bool FULL = false>
__device__ __forceinline__ float4 getprop(const MCXKParam& g, const Shared& sh, uint32_t mediaid, uint32_t idx1d) {
    if (FULL && g.mediaformat > 4) {
        return media_word_prop(g.mediaformat, sh.prop, mediaid);
    }

    float4 p = sh.prop[mediaid & K_MED_MASK];

    /* per-voxel absorption and scattering: g and n stay those of the label */
    if (g.muamus && (mediaid & K_MED_MASK)) {
        float2 m = g.muamus[idx1d];
        p.x = m.x;
        p.y = m.y;
    }

    return p;
}

/* src/mcx_core.cl:1350-1455 */
template <bool FULL>
static __device__ int skipvoid(const MCXKParam& g, Photon& s) {
    int count = 1, idx1d;
    float4& p = s.p;
    float4& v = s.v;
    s.fx = to_short((int)floorf(p.x));
    s.fy = to_short((int)floorf(p.y));
    s.fz = to_short((int)floorf(p.z));
    s.fw = -1;

    while (1) {
        if (inside(g, s.fx, s.fy, s.fz)) {
            idx1d = s.fz * (int)g.dimxy + s.fy * (int)g.dimx + s.fx;

            if (g.media[idx1d] & K_MED_MASK) {
                p.x -= v.x;
                p.y -= v.y;
                p.z -= v.z;
                s.fx = to_short((int)floorf(p.x));
                s.fy = to_short((int)floorf(p.y));
                s.fz = to_short((int)floorf(p.z));
                s.f.y -= g.minaccumtime;
                idx1d = s.fz * (int)g.dimxy + s.fy * (int)g.dimx + s.fx;
                count = 0;

                while (!inside(g, s.fx, s.fy, s.fz) || !(g.media[idx1d] & K_MED_MASK)) {
                    int face;
                    float dist = hitgrid(p, v, s.fx, s.fy, s.fz, &face);
                    s.fw = face;
                    s.f.y += g.minaccumtime * dist;
                    p.x = p.x + dist * v.x;
                    p.y = p.y + dist * v.y;
                    p.z = p.z + dist * v.z;

                    if (face == 0) {
                        s.fx += (v.x > 0.f ? 1 : -1);
                    }

                    if (face == 1) {
                        s.fy += (v.y > 0.f ? 1 : -1);
                    }

                    if (face == 2) {
                        s.fz += (v.z > 0.f ? 1 : -1);
                    }

                    idx1d = s.fz * (int)g.dimxy + s.fy * (int)g.dimx + s.fx;

                    if (count++ > 3) {
                        break;
                    }
                }

                s.f.y = g.voidtime ? s.f.y : 0.f;

                if (!inside(g, s.fx, s.fy, s.fz)) {
                    return -1;   /* entry search failed: the reference would read outside the volume */
                }

                /* n of the entry voxel (a packed media word: its decoded n; the reference
                   indexes the label table with the word, src/mcx_core.cl:1424) */
                float nin;

                if (FULL && g.svmc) {
                    nin = svmc_prop(g, g.prop, g.media[idx1d], (uint32_t)idx1d, p, s.fx, s.fy, s.fz, s).w;
                } else {
                    nin = (FULL && g.mediaformat > 4) ? media_word_prop(g.mediaformat, g.prop, g.media[idx1d]).w :
                          g.prop[g.media[idx1d] & K_MED_MASK].w;
                }

                if (g.isspecular && nin != g.prop[0].w) {
                    p.w *= 1.f - reflectcoeff(v, g.prop[0].w, nin, s.fw);

                    if (p.w > K_EPS) {
                        transmit(v, g.prop[0].w, nin, s.fw);
                    }
                }

                return idx1d;
            }
        }

        if ((p.x < 0.f && v.x <= 0.f) || (p.x >= g.maxidx_x && v.x >= 0.f) || (p.y < 0.f && v.y <= 0.f) ||
                (p.y >= g.maxidx_y && v.y >= 0.f) || (p.z < 0.f && v.z <= 0.f) || (p.z >= g.maxidx_z && v.z >= 0.f)) {
            return -1;
        }

        p.x = p.x + v.x;
        p.y = p.y + v.y;
        p.z = p.z + v.z;
        s.fx = to_short((int)floorf(p.x));
        s.fy = to_short((int)floorf(p.y));
        s.fz = to_short((int)floorf(p.z));
        s.f.y += g.minaccumtime;

        if ((uint32_t)count++ > g.maxvoidstep) {
            return -1;
        }
    }
}

__device__ __forceinline__ uint32_t photon_index(const MCXKParam& g, int idx, float fw) {
    int odd = g.oddphoton;
    return (uint32_t)(idx * (int)g.threadphoton + (idx < odd ? idx : odd) + (int)fw);
}

/* the RNG state at a photon's launch, kept for its detection record in the
   lane's scratch at seedoff (s0 low, s0 high, s1 low, s1 high) when seeds are
   saved: read only at a detection, so it holds no registers across the loop */
__device__ __forceinline__ void seed_put(const MCXKParam& g, float* pp, const Rng& rng) {
    uint32_t* sw = (uint32_t*)(pp + g.seedoff);
    sw[0] = (uint32_t)rng.s0;
    sw[1] = (uint32_t)(rng.s0 >> 32);
    sw[2] = (uint32_t)rng.s1;
    sw[3] = (uint32_t)(rng.s1 >> 32);
}

/* polarised runs keep the Stokes vector {I, Q, U, V} (I kept at 1 after each
   scattering) in the lane's scratch at stokesoff: it is used only at launches,
   scattering events and exits, so it holds no registers across the step loop */
__device__ __forceinline__ float4 stokes_get(const MCXKParam& g, const float* pp) {
    return make_float4(pp[g.stokesoff], pp[g.stokesoff + 1], pp[g.stokesoff + 2], pp[g.stokesoff + 3]);
}

__device__ __forceinline__ void stokes_put(const MCXKParam& g, float* pp, const float4& st) {
    pp[g.stokesoff] = st.x;
    pp[g.stokesoff + 1] = st.y;
    pp[g.stokesoff + 2] = st.z;
    pp[g.stokesoff + 3] = st.w;
}

/* trajectory record (src/mcx_core.cl:929-944): {photon id as bits, x, y, z, w,
   source id}; positions past maxjumpdebug are counted, not stored */
__device__ __forceinline__ void save_traj(const MCXKParam& g, const float4& p, uint32_t id, float srcid,
        const float* stokes = nullptr) {
    const uint32_t pos = atomicAdd(g.jumpdebug, 1u);

    if (pos < g.maxjumpdebug) {
        /* records are 6 floats, 10 with istrajstokes; the Stokes vector follows in
           polarised runs (src/mcx_core.cl:929-959) */
        float* r = g.debugdata + (size_t)pos * g.trajlen;
        r[0] = __uint_as_float(id);
        r[1] = p.x;
        r[2] = p.y;
        r[3] = p.z;
        r[4] = p.w;
        r[5] = (float)(int)srcid;

        if (stokes && g.maxpolmedia && g.istrajstokes) {
            r[6] = stokes[0];
            r[7] = stokes[1];
            r[8] = stokes[2];
            r[9] = stokes[3];
        }
    }
}

/* Stokes vector after a scattering by theta at azimuth phi from direction u to
   u2 in a polarised medium (src/mcx_core.cl:781-836): rotate into the
   scattering plane, apply the Mueller matrix of the angle's table entry, rotate
   into the new meridian plane and normalise to I = 1 */
__device__ __forceinline__ void update_stokes(const MCXKParam& g, float4& st, float theta, float phi, const float4& u,
        const float4& u2, uint32_t mediaid) {
    const float costheta = kcos(theta);
    float sin2phi, cos2phi;
    ksincos(2.f * phi, &sin2phi, &cos2phi);
    float4 s2;
    s2.x = st.x;
    s2.y = st.y * cos2phi + st.z * sin2phi;
    s2.z = -st.y * sin2phi + st.z * cos2phi;
    s2.w = st.w;

    const uint32_t lab = mediaid & K_MED_MASK;
    const float4 m = g.smatrix[K_NANGLES * (lab ? lab - 1 : 0u) + (uint32_t)(theta * (float)K_NANGLES * (K_R_PI - K_EPS))];
    st.x = m.x * s2.x + m.y * s2.y;
    st.y = m.y * s2.x + m.x * s2.y;
    st.z = m.z * s2.z + m.w * s2.w;
    st.w = -m.w * s2.z + m.z * s2.w;

    float temp = (u2.z > -1.f && u2.z < 1.f) ? krsqrt((1.f - costheta * costheta) * (1.f - u2.z * u2.z)) : 0.f;
    float cosi = (temp == 0.f) ? 0.f : (((phi > K_ONE_PI && phi < K_TWO_PI) ? 1.f : -1.f) *
                                        (u2.z * costheta - u.z) * temp);
    cosi = fmaxf(-1.f, fminf(cosi, 1.f));
    const float sini = ksqrt(1.f - cosi * cosi);
    const float cos22 = 2.f * cosi * cosi - 1.f;
    const float sin22 = 2.f * sini * cosi;

    s2.x = st.x;
    s2.y = st.y * cos22 - st.z * sin22;
    s2.z = st.y * sin22 + st.z * cos22;
    s2.w = st.w;

    temp = kdiv(1.f, s2.x);
    st.y = s2.y * temp;
    st.z = s2.z * temp;
    st.w = s2.w * temp;
    st.x = 1.f;
}

/* the photon id the reference writes into launch and scattering records
   (src/mcx_core.cl:2247, :2630): threads past oddphoton add 0 where the
   termination record (:1501) and the photon count add oddphoton */
__device__ __forceinline__ uint32_t traj_index(const MCXKParam& g, int idx, float fw) {
    return (uint32_t)(idx * (int)g.threadphoton + (idx < g.oddphoton ? idx : 0) + (int)fw);
}

__device__ __forceinline__ int gate_of(const MCXKParam& g, float t) {
    int gate = (int)floorf((t - g.twin0) * g.Rtstep);
    return gate > (int)g.maxgate - 1 ? (int)g.maxgate - 1 : gate;
}

/* global deposit into this lane's accumulator replica */
__device__ __forceinline__ void deposit(const MCXKParam& g, size_t repoff, uint32_t addr, float val) {
    atomicAdd(g.field + repoff + addr, (double)val);
}

/* the imaginary half of an RF forward deposit: into the box's second set of
   gates when the voxel and gate fall inside it (same cell rules as
   deposit_voxel), else straight to the grid */
__device__ __forceinline__ void deposit_voxel_im(const MCXKParam& g, const Shared& sh, int ox, int oy, int oz,
        uint32_t idx1d, int tshift, float val) {
    const uint32_t dx = (uint32_t)(ox - (int)g.box_x0), dy = (uint32_t)(oy - (int)g.box_y0);
    const uint32_t dz = (uint32_t)(oz - (int)g.box_z0);
    const uint32_t addr = idx1d + (uint32_t)tshift * g.nvox + g.imoff;

    if (g.box_imgate && dx < g.box_nx && dy < g.box_ny && dz < g.box_nz && (uint32_t)tshift < g.box_ng) {
        BoxCell* cell = sh.box + ((((uint32_t)tshift + g.box_ng) * g.box_nz + dz) * g.box_ny + dy) * g.box_nx + dx;
#if MCX_PARITY
        atomicAdd(cell, (BoxCell)val);
#else
        const float old = atomicAdd(cell, val);

        if (fabsf(old) > K_BOX_MOVE) {
            const float mv = old > 0.f ? K_BOX_MOVE : -K_BOX_MOVE;
            atomicAdd(cell, -mv);
            atomicAdd(g.field + sh.repoff + addr, (double)mv);
        }
#endif
    } else {
        deposit(g, sh.repoff, addr, val);
    }
}

/* src/mcx_core.cl:838-926 */
template <bool SAVEDET, bool FULL>
__device__ void save_detected(const MCXKParam& g, const Shared& sh, const Photon& s, uint32_t isdet) {
    int detid = 0;

    if (isdet == K_OUT_MIN) {
        detid = -1;
    } else {
        for (uint32_t i = 0; i < g.detnum; i++) {
            float4 d = g.detpos[i];

            if ((d.x - s.p.x) * (d.x - s.p.x) + (d.y - s.p.y) * (d.y - s.p.y) + (d.z - s.p.z) * (d.z - s.p.z) <
                    d.w * d.w) {
                detid = (int)i + 1;
                break;
            }
        }
    }

    if (!detid) {
        return;
    }

    uint32_t slot = atomicAdd(g.detcount, 1u);

    if (slot >= g.maxdetphoton) {
        if (g.savedet == K_FILL_MAXDET) {
            atomicSub(g.detcount, 1u);   /* fill-max mode keeps the count at the buffer size (:922-923) */
        }

        return;
    }

    if (g.issaveseed) {
        const uint32_t* sw = (const uint32_t*)(sh.pp + g.seedoff);
        g.detseed[(size_t)slot * 2] = (uint64_t)sw[1] << 32 | sw[0];
        g.detseed[(size_t)slot * 2 + 1] = (uint64_t)sw[3] << 32 | sw[2];
    }

    if (g.detweight) {
        g.detweight[slot] = s.p.w;   /* weight at exit: w0 * exp(-sum mua L) */
    }

    float* rec = g.detrec + (size_t)slot * g.reclen;
    uint32_t k = 0;

    if (g.savedetflag & 0x1) {
        if (g.extrasrclen && g.srcid <= 0) {
            detid |= ((int)sh.pp[g.w0offset - 1]) << 16;
        }

        rec[k++] = (float)detid;
    }

    for (uint32_t i = 0; i < g.partialdata; i++) {
        rec[k++] = sh.pp[i];
    }

    if ((g.savedetflag >> 4) & 1) {
        rec[k++] = s.p.x;
        rec[k++] = s.p.y;
        rec[k++] = s.p.z;
    }

    if ((g.savedetflag >> 5) & 1) {
        rec[k++] = s.v.x;
        rec[k++] = s.v.y;
        rec[k++] = s.v.z;
    }

    if ((g.savedetflag >> 6) & 1) {
        rec[k++] = sh.pp[g.w0offset - 2];
    }

    if (FULL && ((g.savedetflag >> 7) & 1)) {
        /* the Stokes vector at exit (src/mcx_core.cl:915-920) */
        for (uint32_t j = 0; j < 4; j++) {
            rec[k++] = sh.pp[g.stokesoff + j];
        }
    }
}

/* launch weight of a pattern pixel (src/mcx_core.cl:1678-1727): one pattern
   scales the photon's weight; several share a unit-weight photon, each keeping
   its pixel value in the lane scratch (ppl = the scratch past the record
   columns), the pixel index in the w0 slot as bits */
__device__ __forceinline__ void pattern_weights(const MCXKParam& g, float* ppl, float4& p, int pix, float srcw) {
    if (g.npattern <= 1) {
        p.w = srcw * g.srcpattern[pix];
        ppl[4] = p.w;
    } else {
        ppl[2] = __uint_as_float((uint32_t)pix);

        for (uint32_t i = 0; i < g.npattern; i++) {
            ppl[4 + i] = g.srcpattern[(size_t)pix * g.npattern + i];
        }

        p.w = 1.f;
    }
}

/* terminate the current photon and launch the next one (src/mcx_core.cl:1466-2275).
   Returns 0 on a launch, 1 when the lane's photon budget is used up, -1 when
   no launch succeeded within maxvoidstep attempts. */
template <bool SAVEDET, bool SIMPLE, bool FULL>
__device__ int launch_photon(const MCXKParam& g, const Shared& sh, Photon& s, Rng& rng, int idx, uint32_t isdet) {
    float* pp = sh.pp;
    const uint32_t pd = g.partialdata;
    float4& p = s.p;
    float4& v = s.v;
    float4& f = s.f;
    s.w0 = 1.f;
    s.Lmove = -1.f;

    if (SAVEDET && g.savedet == K_FILL_MAXDET && __hip_atomic_load(g.detcount, __ATOMIC_RELAXED,
            __HIP_MEMORY_SCOPE_AGENT) >= g.maxdetphoton) {
        /* fill-max mode (src/mcx_core.cl:1484-1489): the lane stops once the
           detected-photon buffer is full */
        return 1;
    }

    if (fabsf(p.w) >= 0.f) {
        pp[pd] += p.w;

        if (FULL && g.savetraj) {
            save_traj(g, p, photon_index(g, idx, f.w), pp[g.w0offset - 1], pp + g.stokesoff);
        }

        if (!SIMPLE && SAVEDET && g.issaveref > 1 && s.mediaid == 0 && s.idx1d != K_OUT_MIN &&
                s.idx1d != K_OUT_MAX && p.w > 0.f && s.idx1d < g.maxdetphoton) {
            /* exit tally per voxel (src/mcx_core.cl:852-865): the detected-photon
               buffer holds one record per voxel; its first column adds the exit
               weight, and columns 0..maxmedia-1 add each medium's partial path
               times that weight (both writes start at column 0, as in the
               reference).  Atomic here; the reference's plain adds can lose
               updates of concurrent photons. */
            float* rec = g.detrec + (size_t)s.idx1d * g.reclen;
            atomicAdd(rec, p.w);

            for (uint32_t i = 0; i < g.maxmedia; i++) {
                atomicAdd(rec + i, pp[i] * p.w);
            }
        }

        if (!SIMPLE && g.issaveref == 1 && s.mediaid == 0 && s.idx1d != K_OUT_MIN && s.idx1d != K_OUT_MAX &&
                p.w > 0.f) {
            int tshift = (int)g.maxgate - 1;
            int t2 = (int)floorf((f.y - g.twin0) * g.Rtstep);
            tshift = tshift < t2 ? tshift : t2;

            if (g.extrasrclen && g.srcid < 0) {
                tshift += ((int)pp[g.w0offset - 1] - 1) * (int)g.maxgate;
            }

            if (t2 >= 0 && FULL && g.npattern > 1) {
                /* patterns sharing the photon: one interleaved slot per pattern,
                   scaled by its weight (src/mcx_core.cl:1525-1540) */
                const uint32_t base = (s.idx1d + (uint32_t)tshift * g.nvox) * g.npattern;

                for (uint32_t i = 0; i < g.npattern; i++) {
                    if (fabsf(pp[g.w0offset + i]) > 0.f) {
                        deposit(g, sh.repoff, base + i, -(p.w * pp[g.w0offset + i]));
                    }
                }
            } else if (t2 >= 0) {
                deposit(g, sh.repoff, s.idx1d + (uint32_t)tshift * g.nvox, -p.w);
            }
        }

        if (SAVEDET && (isdet & K_DET_MASK) == K_DET_MASK &&
                (s.mediaid == 0 || (FULL && g.svmc && K_SV_CURLABEL(s.sv) == 0)) && g.issaveref < 2) {
            save_detected<SAVEDET, FULL>(g, sh, s, isdet);
        }
    }

    if (SAVEDET) {
        for (uint32_t i = 0; i < pd; i++) {
            pp[i] = 0.f;
        }
    }

    if (g.dynamic) {
        /* the workgroup's photons are handed out one at a time from LDS */
        if (atomicAdd(sh.pool, 1u) >= sh.pool[1]) {
            return 1;
        }
    } else if (f.w >= (float)(g.threadphoton + (idx < g.oddphoton))) {
        return 1;
    }

    if (SIMPLE) {
        /* pencil beam from a non-void voxel, no focus, no extra sources, no replay:
           the general launch below reduces to this (one pass of its loop) */
        if (SAVEDET && g.issaveseed) {
            seed_put(g, pp, rng);
        }

        p = g.srcpos;
        v = g.srcdir;
        f.x = 0.f;
        f.y = 0.f;
        f.z = g.minaccumtime;
        s.idx1d = __float_as_uint(g.srcparam2.z);
        s.mediaid = __float_as_uint(g.srcparam2.w);
        s.fx = to_short((int)floorf(p.x));
        s.fy = to_short((int)floorf(p.y));
        s.fz = to_short((int)floorf(p.z));
        f.w += 1.f;
        s.prop = getprop<FULL>(g, sh, s.mediaid, s.idx1d);
        pp[pd + 1] += p.w;
        s.w0 = p.w;
        pp[pd + 2] = (g.srcnum > 1) ? pp[pd + 2] : p.w;
        v.w = K_EPS;
        s.Lmove = 0.f;
        return 0;
    }

    if (g.seed == K_SEED_FROM_FILE) {
        uint32_t pid = photon_index(g, idx, f.w < 0.f ? 0.f : f.w);
        rng.s0 = g.replayseed[(size_t)pid * 2];
        rng.s1 = g.replayseed[(size_t)pid * 2 + 1];
    }

    if (SAVEDET && g.issaveseed) {
        seed_put(g, pp, rng);
    }

    float4 lpos = g.srcpos, ldir = g.srcdir, lp1 = g.srcparam1, lp2 = g.srcparam2;
    int cursid = 1;

    if (g.extrasrclen && g.srcid != 1) {
        int sid;

        if (g.srcid > 1) {
            sid = g.srcid;
        } else {
            pp[g.w0offset - 1] = (float)((int)(rng.next() * K_JUST_BELOW_ONE * (float)(g.extrasrclen + 1)) + 1);
            sid = (int)pp[g.w0offset - 1];
        }

        cursid = sid;

        if (sid > 1) {
            const float4* es = g.srcdata + (sid - 2) * 4;
            lpos = es[0];
            ldir = es[1];
            lp1 = es[2];
            lp2 = es[3];
        }
    }

    if (g.seed == K_SEED_FROM_FILE && g.srcid >= 1) {
        rng.next();
    }

    float* ppl = pp + pd;

    do {
        p = lpos;
        v = ldir;
        f.x = 0.f;
        f.y = 0.f;
        f.z = g.minaccumtime;
        s.idx1d = __float_as_uint(lp2.z);
        s.mediaid = __float_as_uint(lp2.w);
        s.prop = make_float4(lpos.x, lpos.y, lpos.z, 0.f);

        if (FULL && g.svmc) {
            s.sv = 0;
        }

        if (FULL && g.maxpolmedia) {
            stokes_put(g, pp, g.s0);    /* src/mcx_core.cl:1633-1638 */
        }

        if (!SIMPLE) {
            const int st = (int)g.srctype;
            bool positioned = false;

            if (st == kPlanar || (FULL && (st == kPattern || st == kPattern3D)) || st == kFourier || st == kPencilArray) {
                float rx = rng.next();
                float ry = rng.next();

                if (FULL && st == kPattern3D) {
                    /* a param1.x x param1.y x param1.z box of voxels (src/mcx_core.cl:1661-1667) */
                    const float rz = rng.next();
                    p.x = p.x + rx * lp1.x;
                    p.y = p.y + ry * lp1.y;
                    p.z = p.z + rz * lp1.z;
                    pattern_weights(g, ppl, p, (int)(rz * K_JUST_BELOW_ONE * lp1.z) * (int)(lp1.y) * (int)(lp1.x) +
                                    (int)(ry * K_JUST_BELOW_ONE * lp1.y) * (int)(lp1.x) +
                                    (int)(rx * K_JUST_BELOW_ONE * lp1.x), lpos.w);
                } else {
                    p.x = p.x + rx * lp1.x + ry * lp2.x;
                    p.y = p.y + rx * lp1.y + ry * lp2.y;
                    p.z = p.z + rx * lp1.z + ry * lp2.z;
                }

                if (FULL && st == kPattern) {
                    /* param1.w x param2.w pixels over the parallelogram (src/mcx_core.cl:1678-1693) */
                    pattern_weights(g, ppl, p, (int)(ry * K_JUST_BELOW_ONE * lp2.w) * (int)(lp1.w) +
                                    (int)(rx * K_JUST_BELOW_ONE * lp1.w), lpos.w);
                } else if (st == kFourier) {
                    p.w = lpos.w * (kcos((floorf(lp1.w) * rx + floorf(lp2.w) * ry + lp1.w - floorf(lp1.w)) * K_TWO_PI) *
                                    (1.f - lp2.w + floorf(lp2.w)) + 1.f) * 0.5f;
                } else if (st == kPencilArray) {
                    p.x = lpos.x + kdiv(floorf(rx * lp1.w) * lp1.x, lp1.w - 1.f) + kdiv(floorf(ry * lp2.w) * lp2.x, lp2.w - 1.f);
                    p.y = lpos.y + kdiv(floorf(rx * lp1.w) * lp1.y, lp1.w - 1.f) + kdiv(floorf(ry * lp2.w) * lp2.y, lp2.w - 1.f);
                    p.z = lpos.z + kdiv(floorf(rx * lp1.w) * lp1.z, lp1.w - 1.f) + kdiv(floorf(ry * lp2.w) * lp2.z, lp2.w - 1.f);
                }

                positioned = true;
                s.prop.x = s.prop.x + (lp1.x + lp2.x) * 0.5f;
                s.prop.y = s.prop.y + (lp1.y + lp2.y) * 0.5f;
                s.prop.z = s.prop.z + (lp1.z + lp2.z) * 0.5f;
                s.prop.w = 0.f;
            } else if (FULL && (st == kFourierX || st == kFourierX2D)) {
                /* edges param1 and param1.w dir x param1/|param1|; param2 the pattern
                   (src/mcx_core.cl:1776-1812) */
                const float rx = rng.next();
                const float ry = rng.next();
                float4 v2 = lp1;
                v2.w *= krsqrt(lp1.x * lp1.x + lp1.y * lp1.y + lp1.z * lp1.z);
                v2.x = v2.w * (ldir.y * lp1.z - ldir.z * lp1.y);
                v2.y = v2.w * (ldir.z * lp1.x - ldir.x * lp1.z);
                v2.z = v2.w * (ldir.x * lp1.y - ldir.y * lp1.x);
                p.x = p.x + rx * lp1.x + ry * v2.x;
                p.y = p.y + rx * lp1.y + ry * v2.y;
                p.z = p.z + rx * lp1.z + ry * v2.z;

                if (st == kFourierX2D) {
                    p.w = lpos.w * (ksin((lp2.x * rx + lp2.z) * K_TWO_PI) * ksin((lp2.y * ry + lp2.w) * K_TWO_PI) + 1.f) *
                          0.5f;
                } else {
                    p.w = lpos.w * (kcos((lp2.x * rx + lp2.y * ry + lp2.z) * K_TWO_PI) * (1.f - lp2.w) + 1.f) * 0.5f;
                }

                positioned = true;
                s.prop.x = s.prop.x + (lp1.x + v2.x) * 0.5f;
                s.prop.y = s.prop.y + (lp1.y + v2.y) * 0.5f;
                s.prop.z = s.prop.z + (lp1.z + v2.z) * 0.5f;
                s.prop.w = 0.f;
            } else if (FULL && st == kHyperboloid) {
                /* rays on the one-sheet hyperboloid of a Gaussian beam: waist param1.x at
                   param1.y along the axis, param1.z its length scale; the launch voxel is
                   left as set (src/mcx_core.cl:1820-1852) */
                float sphi, cphi;
                float r = K_TWO_PI * rng.next();
                ksincos(r, &sphi, &cphi);
                r = ksqrt(0.5f * rng.scatlen()) * lp1.x;
                const float hx = kdiv(-lp1.y, lp1.z);
                const float hy = krsqrt(r * r + lp1.z * lp1.z);
                float qx = r * (cphi - hx * sphi), qy = r * (sphi + hx * cphi), qz = 0.f;
                const float dx = -r * sphi * hy, dy = r * cphi * hy, dz = lp1.z * hy;

                if (v.z > -1.f + K_EPS && v.z < 1.f - K_EPS) {
                    float t = 1.f - v.z * v.z;
                    const float stheta = ksqrt(t);
                    t = krsqrt(t);
                    cphi = v.x * t;
                    sphi = v.y * t;
                    const float nx = qx * cphi * v.z - qy * sphi, ny = qx * sphi * v.z + qy * cphi, nz = -qx * stheta;
                    const float vx = dx * cphi * v.z - dy * sphi + dz * cphi * stheta;
                    const float vy = dx * sphi * v.z + dy * cphi + dz * sphi * stheta;
                    const float vz = -dx * stheta + dz * v.z;
                    qx = nx;
                    qy = ny;
                    qz = nz;
                    v.x = vx;
                    v.y = vy;
                    v.z = vz;
                } else {
                    v.z = (v.z > 0.f) ? dz : -dz;
                    v.x = dx;
                    v.y = dy;
                }

                p.x = qx + lpos.x;
                p.y = qy + lpos.y;
                p.z = qz + lpos.z;
                s.prop = make_float4(lpos.x, lpos.y, lpos.z, 0.f);
                s.Lmove = 0.f;
            } else if (st == kDisk || st == kGaussian || st == kRing) {
                float sphi, cphi, phi, r;

                if ((st == kDisk || st == kRing) && (lp1.z > 0.f || lp1.w > 0.f)) {
                    phi = fabsf(lp1.z - lp1.w) * rng.next() + fminf(lp1.z, lp1.w);
                } else {
                    phi = K_TWO_PI * rng.next();
                }

                ksincos(phi, &sphi, &cphi);

                if (st == kDisk || st == kRing) {
                    r = ksqrt(rng.next() * fabsf(lp1.x * lp1.x - lp1.y * lp1.y) + lp1.y * lp1.y);
                } else if (fabsf(ldir.w) < 1e-5f || fabsf(lp1.y) < 1e-5f) {
                    r = ksqrt(-0.5f * klog(rng.next())) * lp1.x;
                } else {
                    r = kdiv(lp1.x * lp1.x * K_ONE_PI, lp1.y);
                    r = ksqrt(-0.5f * klog(rng.next()) * (1.f + kdiv(ldir.w * ldir.w, r * r))) * lp1.x;
                }

                if (v.z > -1.f + K_EPS && v.z < 1.f - K_EPS) {
                    float tmp0 = 1.f - v.z * v.z;
                    float tmp1 = r * krsqrt(tmp0);
                    p.x = p.x + tmp1 * (v.x * v.z * cphi - v.y * sphi);
                    p.y = p.y + tmp1 * (v.y * v.z * cphi + v.x * sphi);
                    p.z = p.z - tmp1 * tmp0 * cphi;
                } else {
                    p.x += r * cphi;
                    p.y += r * sphi;
                }

                positioned = true;
            } else if (st == kCone || st == kIsotropic || st == kArcsine) {
                float ang, stheta, ctheta, sphi, cphi;
                ang = K_TWO_PI * rng.next();
                ksincos(ang, &sphi, &cphi);

                if (st == kCone) {
                    do {
                        ang = (lp1.y > 0) ? K_TWO_PI * rng.next() : kacos(2.f * rng.next() - 1.f);
                    } while (ang > lp1.x);
                } else if (st == kIsotropic) {
                    ang = kacos(2.f * rng.next() - 1.f);
                } else {
                    ang = K_ONE_PI * rng.next();
                }

                ksincos(ang, &stheta, &ctheta);
                rotatevector(v, stheta, ctheta, sphi, cphi);
                s.Lmove = 0.f;
            } else if (st == kZGaussian) {
                float ang, stheta, ctheta, sphi, cphi;
                ang = K_TWO_PI * rng.next();
                ksincos(ang, &sphi, &cphi);
                float u1 = rng.next();
                float u2 = rng.next();
                ang = ksqrt(-2.f * klog(u1)) * (1.f - 2.f * u2) * lp1.x;
                ksincos(ang, &stheta, &ctheta);
                rotatevector(v, stheta, ctheta, sphi, cphi);
                s.Lmove = 0.f;
            } else if (st == kLine || st == kSlit) {
                float r_l = rng.next();
                p.x = p.x + r_l * lp1.x;
                p.y = p.y + r_l * lp1.y;
                p.z = p.z + r_l * lp1.z;

                if (st == kLine) {
                    float sphi_l, cphi_l;
                    r_l = krsqrt(lp1.x * lp1.x + lp1.y * lp1.y + lp1.z * lp1.z);

                    if (lp2.x > 0.f) {
                        float ax = lp1.x * r_l, ay = lp1.y * r_l, az = lp1.z * r_l;
                        float vdotaxis = v.x * ax + v.y * ay + v.z * az;
                        v.x = v.x - vdotaxis * ax;
                        v.y = v.y - vdotaxis * ay;
                        v.z = v.z - vdotaxis * az;
                        float vnorm = krsqrt(v.x * v.x + v.y * v.y + v.z * v.z);
                        v.x = v.x * vnorm;
                        v.y = v.y * vnorm;
                        v.z = v.z * vnorm;
                        r_l = lp2.x * (2.f * rng.next() - 1.f);
                        ksincos(r_l, &sphi_l, &cphi_l);
                        float cx = ay * v.z - az * v.y;
                        float cy = az * v.x - ax * v.z;
                        float cz = ax * v.y - ay * v.x;
                        v.x = v.x * cphi_l + cx * sphi_l;
                        v.y = v.y * cphi_l + cy * sphi_l;
                        v.z = v.z * cphi_l + cz * sphi_l;
                    } else {
                        v.x = lp1.x * r_l;
                        v.y = lp1.y * r_l;
                        v.z = lp1.z * r_l;
                        r_l = K_TWO_PI * rng.next();
                        ksincos(r_l, &sphi_l, &cphi_l);
                        rotatevector(v, 1.f, 0.f, sphi_l, cphi_l);
                    }
                } else if (lp2.x > 0.f || lp2.y > 0.f) {
                    float sphi_s, cphi_s;
                    r_l = K_TWO_PI * rng.next();
                    ksincos(r_l, &sphi_s, &cphi_s);
                    r_l = ksqrt(2.f * rng.scatlen());
                    cphi_s *= lp2.x * r_l;
                    sphi_s *= lp2.y * r_l;
                    sphi_s *= krsqrt(lp1.x * lp1.x + lp1.y * lp1.y + lp1.z * lp1.z);
                    float4 c = make_float4(lp1.y * v.z - lp1.z * v.y, lp1.z * v.x - lp1.x * v.z, lp1.x * v.y - lp1.y * v.x, 0.f);
                    cphi_s *= krsqrt(c.x * c.x + c.y * c.y + c.z * c.z);
                    v.x += cphi_s * c.x + sphi_s * lp1.x;
                    v.y += cphi_s * c.y + sphi_s * lp1.y;
                    v.z += cphi_s * c.z + sphi_s * lp1.z;
                    r_l = krsqrt(v.x * v.x + v.y * v.y + v.z * v.z);
                    v.x *= r_l;
                    v.y *= r_l;
                    v.z *= r_l;
                }

                positioned = true;
                s.prop.x = lpos.x + lp1.x * 0.5f;
                s.prop.y = lpos.y + lp1.y * 0.5f;
                s.prop.z = lpos.z + lp1.z * 0.5f;
                s.prop.w = 0.f;
                s.Lmove = -1.f;
            }

            if (positioned) {
                s.idx1d = (uint32_t)((int)floorf(p.z) * (int)g.dimxy + (int)floorf(p.y) * (int)g.dimx + (int)floorf(p.x));

                if (p.x < 0.f || p.y < 0.f || p.z < 0.f || p.x >= g.maxidx_x || p.y >= g.maxidx_y || p.z >= g.maxidx_z) {
                    s.mediaid = 0;
                } else {
                    s.mediaid = g.media[s.idx1d];
                }
            }
        }

        if (fabsf(p.w) <= g.minenergy) {
            continue;   /* re-tests the loop condition, as the reference does */
        }

        if (FULL && g.nangle) {
            /* sampled launch polar angle (src/mcx_core.cl:2098-2125): theta/pi at nangle
               equal CDF steps, taken as steps (srcdir.w > 0) or interpolated; uniform
               azimuth about srcdir (used as is when 0 <= srcdir.w < 1.5) */
            float ang, stheta, ctheta, sphi, cphi;

            if (ldir.w > 0.f) {
                ang = fminf(rng.next() * (float)g.nangle, (float)g.nangle - K_EPS);
                cphi = g.angleinvcdf[(int)ang];
            } else {
                ang = fminf(rng.next() * (float)(g.nangle - 1), (float)(g.nangle - 1) - K_EPS);
                sphi = ang - (float)((int)ang);
                const uint32_t i0 = ((uint32_t)ang >= g.nangle - 1) ? g.nangle - 1 : (uint32_t)(int)ang;
                const uint32_t i1 = ((uint32_t)ang + 1 >= g.nangle - 1) ? g.nangle - 1 : (uint32_t)((int)ang + 1);
                cphi = (1.f - sphi) * g.angleinvcdf[i0] + sphi * g.angleinvcdf[i1];
            }

            cphi *= K_ONE_PI;
            ksincos(cphi, &stheta, &ctheta);
            ang = K_TWO_PI * rng.next();
            ksincos(ang, &sphi, &cphi);

            if (ldir.w < 1.5f && ldir.w >= 0.f) {
                v = ldir;
            }

            rotatevector(v, stheta, ctheta, sphi, cphi);
        } else if (s.Lmove < 0.f) {
            if (ldir.w != ldir.w) {
                float ang, stheta, ctheta, sphi, cphi;
                ang = K_TWO_PI * rng.next();
                ksincos(ang, &sphi, &cphi);
                ang = kacos(2.f * rng.next() - 1.f);
                ksincos(ang, &stheta, &ctheta);
                rotatevector(v, stheta, ctheta, sphi, cphi);
            } else if (ldir.w < 0.f && isinf(ldir.w)) {
                float ang, stheta, ctheta, sphi, cphi;
                ang = K_TWO_PI * rng.next();
                ksincos(ang, &sphi, &cphi);
                stheta = ksqrt(rng.next());
                ctheta = ksqrt(1.f - stheta * stheta);
                rotatevector(v, stheta, ctheta, sphi, cphi);
            } else if (ldir.w != 0.f) {
                float Rn2 = (float)((ldir.w > 0.f) - (ldir.w < 0.f));
                s.prop.x += ldir.w * v.x;
                s.prop.y += ldir.w * v.y;
                s.prop.z += ldir.w * v.z;
                v.x = Rn2 * (s.prop.x - p.x);
                v.y = Rn2 * (s.prop.y - p.y);
                v.z = Rn2 * (s.prop.z - p.z);
                Rn2 = krsqrt(v.x * v.x + v.y * v.y + v.z * v.z);
                v.x *= Rn2;
                v.y *= Rn2;
                v.z *= Rn2;
            }
        }

        /* adjoint outputs: a detector source launches from a disk of the detector's
           radius (srcparam1.x) across its direction (src/mcx_core.cl:2159-2184) */
        if (!SIMPLE && g.adjsrc && cursid > (int)g.adjsrc) {
            float sphi, cphi;
            ksincos(K_TWO_PI * rng.next(), &sphi, &cphi);
            const float r = ksqrt(rng.next()) * lp1.x;

            if (v.z > -1.f + K_EPS && v.z < 1.f - K_EPS) {
                const float t0 = 1.f - v.z * v.z;
                const float t1 = r * krsqrt(t0);
                p.x += t1 * (v.x * v.z * cphi - v.y * sphi);
                p.y += t1 * (v.y * v.z * cphi + v.x * sphi);
                p.z -= t1 * t0 * cphi;
            } else {
                p.x += r * cphi;
                p.y += r * sphi;
            }

            s.idx1d = (uint32_t)((int)floorf(p.z) * (int)g.dimxy + (int)floorf(p.y) * (int)g.dimx + (int)floorf(p.x));

            if (p.x < 0.f || p.y < 0.f || p.z < 0.f || p.x >= g.maxidx_x || p.y >= g.maxidx_y || p.z >= g.maxidx_z) {
                s.mediaid = 0;
            } else {
                s.mediaid = g.media[s.idx1d];
            }
        }

        if ((s.mediaid & K_MED_MASK) == 0) {
            int id = skipvoid<FULL>(g, s);

            if (id >= 0) {
                s.idx1d = (uint32_t)id;
                s.mediaid = g.media[s.idx1d];
            }
        }

        s.fx = to_short((int)floorf(p.x));
        s.fy = to_short((int)floorf(p.y));
        s.fz = to_short((int)floorf(p.z));
        s.w0 += 1.f;

        if (s.w0 > (float)g.maxvoidstep) {
            return -1;
        }
    } while ((s.mediaid & K_MED_MASK) == 0 || fabsf(p.w) <= g.minenergy);

    f.w += 1.f;
    s.prop = (FULL && g.svmc) ? svmc_prop(g, sh.prop, s.mediaid, s.idx1d, p, s.fx, s.fy, s.fz, s) :
             getprop<FULL>(g, sh, s.mediaid, s.idx1d);

    if (FULL && g.savetraj) {
        save_traj(g, p, traj_index(g, idx, f.w), ppl[3], pp + g.stokesoff);   /* the source-id slot */
    }

    ppl[1] += p.w;
    s.w0 = p.w;
    ppl[2] = (g.srcnum > 1) ? ppl[2] : p.w;
    v.w = K_EPS;
    s.Lmove = 0.f;

    if (FULL && g.imoff && g.seed == K_SEED_FROM_FILE) {
        /* RF replay: the phase omega * tof of this photon (src/mcx_core.cl:2257-2263,
           which reads the next photon's time of flight; here the photon's own) */
        float rs, rc;
        ksincos(g.omega * g.replaytof[photon_index(g, idx, f.w) - 1], &rs, &rc);
        ppl[4 + g.srcnum] = rc;
        ppl[5 + g.srcnum] = rs;
    }

    return 0;
}

/* Transport this lane's photons until its share (fixed map) or its
   workgroup's pool (run-time hand-out) is exhausted.  The photon state starts
   fresh; the RNG, the LDS tables and the lane's scratch come in, and the
   lane's pending grid deposits are flushed before returning. */
template <bool SAVEDET, bool REFLECT, bool SIMPLE, bool FULL>
__device__ __forceinline__ bool transport_lane(const MCXKParam& g, const Shared& sh, Rng& rng, const int idx) {
    Photon s;
    s.p = make_float4(0.f, 0.f, 0.f, __uint_as_float(0x7FC00000u));
    s.v = make_float4(0.f, 0.f, 0.f, 0.f);
    s.f = make_float4(0.f, 0.f, 0.f, 0.f);
    s.prop = make_float4(0.f, 0.f, 0.f, 0.f);
    s.fx = s.fy = s.fz = 0;
    s.fw = -1;
    s.idx1d = 0;
    s.mediaid = __float_as_uint(g.srcparam2.w);
    s.w0 = 0.f;
    s.Lmove = 0.f;

    uint32_t isdet = 0, mediaidold = 0, idx1dold = 0, launchflag = 0;
    float pathlen = 0.f, n1 = 0.f;
    bool relaunch = true;
    /* coordinates of the voxel idx1dold names.  After a reflection idx1d is restored
       while the integer position keeps the reflected face's index for one step, so the
       previous coordinates are kept across that step. */
    int ox = 0, oy = 0, oz = 0;
    bool keepold = false;
    DepCache dcache;
    dcache.a0 = dcache.a1 = 0xFFFFFFFFu;
    dcache.v0 = dcache.v1 = 0.0;
    /* RF forward: the complex photon weight; its value at the last voxel exit is
       read only at exits and sits in the lane scratch at rfoff (real, imaginary) */
    struct {
        float wre, wim;
    } rf = {0.f, 0.f};
    /* split voxels: test the plane on this step (cleared once the photon crossed
       it, set by a scattering event, a voxel change or some relaunches), and
       whether the step ended on it (src/mcx_core.cl:2342-2343) */
    bool testint = false, hitintf = false;
    s.sv = 0;

    if (FULL && g.svmc) {
        s.sn = sh.pp + g.svoff;
        s.sn[0] = s.sn[1] = s.sn[2] = s.sn[3] = 0.f;
    }
    /* move-only trajectories: the lane's run ends when a photon moves into the
       background without being relaunched */
    bool moveend = false;

    for (;;) {
        if (relaunch) {
            if (launch_photon<SAVEDET, SIMPLE, FULL>(g, sh, s, rng, idx, launchflag)) {
                break;
            }

            isdet = s.mediaid & K_DET_MASK;
            s.mediaid &= K_MED_MASK;
            relaunch = false;
            keepold = false;

            if (FULL && g.rfforward) {
                rf.wre = sh.pp[g.rfoff] = s.p.w;
                rf.wim = sh.pp[g.rfoff + 1] = 0.f;
            }
        }

        float4& p = s.p;
        float4& v = s.v;
        float4& f = s.f;

        if (f.x <= 0.f) {
            f.x = rng.scatlen();

            if (v.w != K_EPS) {
                float cphi = 1.f, sphi = 0.f, theta = 0.f, stheta, ctheta, tmp0 = 0.f;

                if (FULL && g.maxpolmedia && !g.is2d) {
                    /* polarised media (src/mcx_core.cl:2454-2468): (theta, phi) by rejection
                       against I(theta, phi) = S11 I + S12 (Q cos 2phi + U sin 2phi) of the
                       medium's table, the forward entry as the bound */
                    const uint32_t lab = s.mediaid & K_MED_MASK;
                    const uint32_t ipol = K_NANGLES * (lab ? lab - 1 : 0u);   /* scattering needs mus > 0: label >= 1 */
                    const float4 st = stokes_get(g, sh.pp);
                    float ii, sin2phi, cos2phi;
#if MCX_PARITY
                    float i0;

                    do {
                        theta = kacos(2.f * rng.next() - 1.f);
                        tmp0 = K_TWO_PI * rng.next();
                        ksincos(2.f * tmp0, &sin2phi, &cos2phi);
                        const float4 m0 = g.smatrix[ipol];
                        i0 = m0.x * st.x + m0.y * (st.y * cos2phi + st.z * sin2phi);
                        const float4 mt = g.smatrix[ipol + (uint32_t)(theta * (float)K_NANGLES * (K_R_PI - K_EPS))];
                        ii = mt.x * st.x + mt.y * (st.y * cos2phi + st.z * sin2phi);
                    } while (rng.next() * i0 >= ii);

                    ksincos(theta, &stheta, &ctheta);
#else
                    /* the same distribution in fewer draws: theta from its marginal, S11 of
                       its table bin times sin(theta) (a CDF over the bins, uniform in
                       cos(theta) inside one), then phi given theta by rejection against
                       S11 I + |S12| |(Q, U)|, which accepts at least half the draws (the
                       joint rejection above takes ~17 rounds for 0.3 um spheres) */
                    const float* cdf = g.polcdf + (K_NANGLES + 1) * (1u + (lab ? lab - 1 : 0u));
                    const float u = rng.next();
                    uint32_t lo = 0, hi = K_NANGLES;

                    while (hi - lo > 1) {
                        const uint32_t mid = (lo + hi) >> 1;

                        if (cdf[mid] <= u) {
                            lo = mid;
                        } else {
                            hi = mid;
                        }
                    }

                    const float ce0 = g.polcdf[lo], ce1 = g.polcdf[lo + 1];
                    ctheta = ce1 + (ce0 - ce1) * rng.next();
                    theta = kacos(ctheta);
                    stheta = ksin_acos(ctheta);
                    const float4 mt = g.smatrix[ipol + lo];
                    const float bound = mt.x * st.x +
                                        fabsf(mt.y) * ksqrt(st.y * st.y + st.z * st.z);

                    do {
                        tmp0 = K_TWO_PI * rng.next();
                        ksincos(2.f * tmp0, &sin2phi, &cos2phi);
                        ii = mt.x * st.x + mt.y * (st.y * cos2phi + st.z * sin2phi);
                    } while (rng.next() * bound >= ii);
#endif
                    ksincos(tmp0, &sphi, &cphi);
                } else {
                if (SIMPLE || !g.is2d) {
                    tmp0 = K_TWO_PI * rng.next();
                    ksincos(tmp0, &sphi, &cphi);
                }

                if (!SIMPLE && g.nphase > 2) {
                    tmp0 = rng.next() * (float)(g.nphase - 1);
                    theta = tmp0 - (float)((int)tmp0);
                    uint32_t i0 = (uint32_t)tmp0 >= g.nphase ? g.nphase - 1 : (uint32_t)(int)tmp0;
                    uint32_t i1 = (uint32_t)tmp0 + 1 >= g.nphase ? g.nphase - 1 : (uint32_t)((int)tmp0 + 1);
                    tmp0 = (1.f - theta) * sh.invcdf[i0] + theta * sh.invcdf[i1];
                    stheta = ksin_acos(tmp0);
                    ctheta = tmp0;

                    if (FULL && g.maxpolmedia) {
                        theta = kacos(tmp0);    /* the angle the Stokes update takes (2D polarised runs) */
                    }
                } else {
                    tmp0 = (v.w > (float)g.gscatter) ? 0.f : s.prop.z;

                    if (fabsf(tmp0) > K_EPS) {
                        tmp0 = kdiv(1.f - s.prop.z * s.prop.z, 1.f - s.prop.z + 2.f * s.prop.z * rng.next());
                        tmp0 *= tmp0;
                        tmp0 = kdiv(1.f + s.prop.z * s.prop.z - tmp0, 2.f * s.prop.z);
                        tmp0 = fmaxf(-1.f, fminf(1.f, tmp0));
                        stheta = ksin_acos(tmp0);
                        ctheta = tmp0;

                        if (FULL && g.maxpolmedia) {
                            theta = kacos(tmp0);
                        }
                    } else {
#if MCX_PARITY
                        theta = kacos(2.f * rng.next() - 1.f);
                        ksincos(theta, &stheta, &ctheta);
#else
                        ctheta = 2.f * rng.next() - 1.f;
                        stheta = ksin_acos(ctheta);

                        if (FULL && g.maxpolmedia) {
                            theta = kacos(ctheta);
                        }
#endif
                    }
                }
                }

                /* per-medium tallies: a split voxel's current side (src/mcx_core.cl:2502-2540) */
                const uint32_t tlab = (FULL && g.svmc) ? K_SV_CURLABEL(s.sv) : (s.mediaid & K_MED_MASK);

                if (SAVEDET && tlab > 0) {
                    if ((g.savedetflag >> 1) & 1) {
                        uint32_t* cnt = (uint32_t*)sh.pp;
                        cnt[tlab - 1]++;
                    }

                    if ((g.savedetflag >> 3) & 1) {
                        sh.pp[g.maxmedia * (((g.savedetflag >> 1) & 1) + ((g.savedetflag >> 2) & 1)) + tlab - 1] +=
                            1.f - ctheta;
                    }
                }

                const float4 olddir = v;

                if (!SIMPLE && g.is2d) {
                    rotatevector2d(v, (rng.next() > 0.5f ? stheta : -stheta), ctheta, (int)g.is2d);
                } else {
                    rotatevector(v, stheta, ctheta, sphi, cphi);
                }

                v.w += 1.f;

                if (FULL && g.maxpolmedia) {
                    /* :2550-2565; in 2D the reference samples by HG with the medium's g and
                       passes the sampled cosine (tmp0) where the azimuth goes */
                    float4 st = stokes_get(g, sh.pp);
                    update_stokes(g, st, theta, tmp0, olddir, v, s.mediaid);
                    stokes_put(g, sh.pp, st);
                }

                if (!SIMPLE && (g.outputtype == kWP || g.outputtype == kDCS ||
                                (FULL && (g.outputtype == kWPTOF ||
                                          (g.seed == K_SEED_FROM_FILE && g.outputtype == kRFmus))))) {
                    /* replay deposits at each scattering site (src/mcx_core.cl:2568-2623):
                       w (WP), (1 - cos theta) w (DCS), tof w (WPTOF), w e^{i omega tof} (RFmus) */
                    uint32_t pid = photon_index(g, idx, f.w) - 1;
                    float im = 0.f;
                    tmp0 = (g.outputtype == kDCS) ? (1.f - ctheta) : 1.f;

                    if (FULL && g.outputtype == kRFmus) {
                        const float rw = g.replayweight[pid];
                        tmp0 = rw * sh.pp[g.w0offset + g.srcnum];
                        im = rw * sh.pp[g.w0offset + g.srcnum + 1];
                    } else if (FULL && g.outputtype == kWPTOF) {
                        tmp0 = g.replaytof[pid];
                        tmp0 *= g.replayweight[pid];
                    } else {
                        tmp0 *= g.replayweight[pid];
                    }

                    int tshift = gate_of(g, g.replaytof[pid]) +
                                 ((g.replaydet == -1) ? (((g.replaydetid[pid] & 0xFFFF) - 1) * (int)g.maxgate) : 0);

                    if (g.extrasrclen && g.srcid < 0) {
                        tshift += ((int)sh.pp[g.w0offset - 1] - 1) * ((g.replaydet == -1) ? (int)g.detnum : 1) *
                                  (int)g.maxgate;
                    }

                    deposit(g, sh.repoff, s.idx1d + (uint32_t)tshift * g.nvox, tmp0);

                    if (FULL && g.outputtype == kRFmus) {
                        deposit(g, sh.repoff, s.idx1d + (uint32_t)tshift * g.nvox + g.imoff, im);
                    }
                }

                if (FULL && g.savetraj) {
                    save_traj(g, p, traj_index(g, idx, f.w), sh.pp[g.w0offset - 1], sh.pp + g.stokesoff);
                }
            }

            v.w = (float)(int)v.w;

            if (FULL && g.svmc) {
                testint = true;
            }
        }

        n1 = s.prop.w;
        /* split voxels: the side is taken again from the position at every step
           (src/mcx_core.cl:2656-2672) */
        s.prop = (FULL && g.svmc) ? svmc_prop(g, sh.prop, s.mediaid, s.idx1d, p, s.fx, s.fy, s.fz, s) :
                 getprop<FULL>(g, sh, s.mediaid, s.idx1d);
        int face;
        f.z = hitgrid(p, v, s.fx, s.fy, s.fz, &face);
        s.fw = face;
        float gfac = (v.w + 1.f > (float)g.gscatter ? (1.f - s.prop.z) : 1.f);
        float slen = f.z * s.prop.y * gfac;
        slen = fminf(slen, f.x);
        f.z = kdiv(slen, s.prop.y * gfac);

        if (FULL && g.svmc) {
            /* the step ends on the voxel's plane when it crosses it (src/mcx_core.cl:2684-2701) */
            hitintf = false;

            if (K_SV_ISSPLIT(s.sv) && testint) {
                float len = f.z;
                hitintf = svmc_ray_plane(s, v, gfac, &len, &slen);
                f.z = len;
            }
        }
        if (!keepold) {
            ox = s.fx;
            oy = s.fy;
            oz = s.fz;
        }

        keepold = false;
        pathlen += f.z;
        p.x = p.x + f.z * v.x;
        p.y = p.y + f.z * v.y;
        p.z = p.z + f.z * v.z;

        const bool stay = (slen == f.x) || (FULL && hitintf);

        if (face == 0) {
            s.fx += stay ? 0 : (v.x > 0.f ? 1 : -1);
        }

        if (face == 1) {
            s.fy += stay ? 0 : (v.y > 0.f ? 1 : -1);
        }

        if (face == 2) {
            s.fz += stay ? 0 : (v.z > 0.f ? 1 : -1);
        }

        if (FULL && g.rfforward) {
            /* RF forward (src/mcx_core.cl:2750-2760): w <- w exp(-(mua + i omega n / c0) ds),
               p.w its magnitude */
            const float att = kexp(-s.prop.x * f.z);
            float rs, rc;
            ksincos(g.omega * s.prop.w * g.oneoverc0 * f.z, &rs, &rc);
            const float tre = att * (rf.wre * rc + rf.wim * rs);
            const float tim = att * (-rf.wre * rs + rf.wim * rc);
            rf.wre = tre;
            rf.wim = tim;
            p.w = ksqrt(rf.wre * rf.wre + rf.wim * rf.wim);
        } else {
            p.w *= kexp(-s.prop.x * f.z);
        }

        f.x -= slen;
        f.y += f.z * s.prop.w * g.oneoverc0;

        if (SAVEDET && ((g.savedetflag >> 2) & 1)) {
            const uint32_t plab = (FULL && g.svmc) ? K_SV_CURLABEL(s.sv) : (s.mediaid & K_MED_MASK);

            if (plab > 0) {
                sh.pp[g.maxmedia * ((g.savedetflag >> 1) & 1) + plab - 1] += f.z;
            }
        }

        mediaidold = s.mediaid | isdet;
        idx1dold = s.idx1d;
        s.idx1d = (uint32_t)(s.fz * (int)g.dimxy + s.fy * (int)g.dimx + s.fx);

        if (!inside(g, s.fx, s.fy, s.fz)) {
            s.mediaid = 0;
            s.idx1d = (s.fx < 0 || s.fy < 0 || s.fz < 0) ? K_OUT_MIN : K_OUT_MAX;
            isdet = g.bc[(s.idx1d == K_OUT_MAX) * 3 + face];
            isdet = ((isdet & 0xF) == 0) ? (g.doreflect ? (uint32_t)kReflect : (uint32_t)kAbsorb) : isdet;
        } else {
            s.mediaid = g.media[s.idx1d];
            isdet = s.mediaid & K_DET_MASK;
            s.mediaid &= K_MED_MASK;
        }

        if ((s.idx1d != idx1dold || (FULL && hitintf)) && idx1dold < g.nvox && mediaidold) {
            if (g.save2pt && f.y >= g.twin0 && f.y < g.twin1) {
                float weight = 0.f, weight_im = 0.f;
                int tshift = gate_of(g, f.y);

                if (FULL && g.rfforward) {
                    /* RF forward (src/mcx_core.cl:2833-2839): (w0 - w) / (mua + i omega n / c0) */
                    const float w0re = sh.pp[g.rfoff], w0im = sh.pp[g.rfoff + 1];
                    const float dre = w0re - rf.wre, dim = w0im - rf.wim;
                    const float aim = g.omega * s.prop.w * g.oneoverc0;
                    const float amag2 = s.prop.x * s.prop.x + aim * aim;
                    weight = (amag2 < K_EPS) ? (w0re * pathlen) : kdiv(dre * s.prop.x + dim * aim, amag2);
                    weight_im = (amag2 < K_EPS) ? (w0im * pathlen) : kdiv(dim * s.prop.x - dre * aim, amag2);
                } else if (g.outputtype == kEnergy) {
                    weight = s.w0 - p.w;
                } else if (g.outputtype == kFluence || g.outputtype == kFlux) {
                    weight = (s.prop.x < K_EPS) ? (s.w0 * pathlen) : kdiv(s.w0 - p.w, s.prop.x);
                } else if (!SIMPLE && g.seed == K_SEED_FROM_FILE) {
                    if (g.outputtype == kJacobian || (FULL && (g.outputtype == kRF || g.outputtype == kWLTOF))) {
                        /* replay path deposits (src/mcx_core.cl:2845-2858): w L (J),
                           -w L e^{i omega tof} (RF), tof w L (WLTOF) */
                        uint32_t pid = photon_index(g, idx, f.w) - 1;
                        weight = g.replayweight[pid] * pathlen;
                        tshift = gate_of(g, g.replaytof[pid]) +
                                 ((g.replaydet == -1) ? (((g.replaydetid[pid] & 0xFFFF) - 1) * (int)g.maxgate) : 0);

                        if (FULL && g.outputtype == kRF) {
                            weight_im = -weight * sh.pp[g.w0offset + g.srcnum + 1];
                            weight = -weight * sh.pp[g.w0offset + g.srcnum];
                        } else if (FULL && g.outputtype == kWLTOF) {
                            weight = weight * g.replaytof[pid];
                        }
                    }
                } else if (g.outputtype == kL) {
                    weight = s.w0 * pathlen;
                }

                if (!SIMPLE && g.extrasrclen && g.srcid < 0) {
                    tshift += ((int)sh.pp[g.w0offset - 1] - 1) * ((g.replaydet == -1) ? (int)g.detnum : 1) *
                              (int)g.maxgate;
                }

                if (FULL && g.npattern > 1) {
                    /* several patterns share the photon: one interleaved slot per pattern
                       (src/mcx_core.cl:2902-2914), straight to the grid */
                    const uint32_t base = (idx1dold + (uint32_t)tshift * g.nvox) * g.npattern;

                    for (uint32_t i = 0; i < g.npattern; i++) {
                        const float pw = sh.pp[g.w0offset + i];

                        if (fabsf(pw) > 0.f && fabsf(weight) > 0.f) {
                            deposit(g, sh.repoff, base + i, weight * pw);
                        }
                    }
                } else if (fabsf(weight) > 0.f) {
                    deposit_voxel(g, sh, dcache, ox, oy, oz, idx1dold, tshift, weight);

                    /* a pattern source's deposit carries no imaginary term in the
                       reference (src/mcx_core.cl:2902-2914): its RF forward field
                       stays real */
                    if (FULL && g.rfforward && g.srctype != kPattern && g.srctype != kPattern3D) {
                        deposit_voxel_im(g, sh, ox, oy, oz, idx1dold, tshift, weight_im);
                    }
                }

                if (FULL && g.outputtype == kRF && g.seed == K_SEED_FROM_FILE && g.omega > 0.f &&
                        fabsf(weight_im) > 0.f) {
                    deposit(g, sh.repoff, idx1dold + (uint32_t)tshift * g.nvox + g.imoff, weight_im);
                }
            }

            s.w0 = p.w;

            if (FULL && g.rfforward) {
                sh.pp[g.rfoff] = rf.wre;
                sh.pp[g.rfoff + 1] = rf.wim;
            }

            pathlen = 0.f;
        } else {
            s.mediaid = mediaidold;
        }

        if (FULL && g.svmc) {
            /* split voxels (src/mcx_core.cl:2937-2949): a new voxel brings its own
               plane and side; a plane crossing moves the photon to the other side */
            if (s.idx1d != idx1dold) {
                s.prop = svmc_prop(g, sh.prop, s.mediaid, s.idx1d, p, s.fx, s.fy, s.fz, s);
                testint = true;
            } else if (hitintf) {
                s.sn[0] = -s.sn[0];
                s.sn[1] = -s.sn[1];
                s.sn[2] = -s.sn[2];
                s.sn[3] = -s.sn[3];
                s.sv ^= 1u << 17;
                testint = false;
            }
        }

        /* split voxels also end the photon on entering label 0 through a voxel face
           or the plane, unless it can reflect there (src/mcx_core.cl:2959-2966) */
        const bool svexit = FULL && g.svmc && (s.idx1d != idx1dold || hitintf) && !K_SV_ISUPPER(s.sv) &&
                            !K_SV_LOWER(s.sv) && (!g.doreflect || n1 == sh.prop[0].w);

        if ((s.mediaid == 0 && ((isdet & 0xF) == kAbsorb || (isdet & 0xF) == kCyclic ||
                                ((isdet & 0xF) == kReflect && n1 == sh.prop[0].w))) || svexit || !(f.y <= g.twin1)) {
            if (isdet == kCyclic) {
                if (face == 0) {
                    p.x = nextafter_ref(rintf((s.idx1d == K_OUT_MIN) ? g.maxidx_x : 0.f), (v.x > 0.f) - (v.x < 0.f));
                    s.fx = to_short((int)floorf(p.x));
                }

                if (face == 1) {
                    p.y = nextafter_ref(rintf((s.idx1d == K_OUT_MIN) ? g.maxidx_y : 0.f), (v.y > 0.f) - (v.y < 0.f));
                    s.fy = to_short((int)floorf(p.y));
                }

                if (face == 2) {
                    p.z = nextafter_ref(rintf((s.idx1d == K_OUT_MIN) ? g.maxidx_z : 0.f), (v.z > 0.f) - (v.z < 0.f));
                    s.fz = to_short((int)floorf(p.z));
                }

                if (inside(g, s.fx, s.fy, s.fz)) {
                    s.idx1d = (uint32_t)(s.fz * (int)g.dimxy + s.fy * (int)g.dimx + s.fx);
                    s.mediaid = g.media[s.idx1d];
                    isdet = s.mediaid & K_DET_MASK;
                    s.mediaid &= K_MED_MASK;
                    continue;
                }
            }

            launchflag = (((s.idx1d == K_OUT_MAX && g.bc[9 + face]) || (s.idx1d == K_OUT_MIN && g.bc[6 + face])) ?
                          K_OUT_MIN : (mediaidold & K_DET_MASK));
            relaunch = true;
            testint = testint || (FULL && g.svmc);
            continue;
        }

        if (fabsf(p.w) < g.minenergy) {
            if (rng.next() * K_ROULETTE <= 1.f) {
                p.w *= K_ROULETTE;

                if (FULL && g.rfforward) {
                    rf.wre *= K_ROULETTE;
                    rf.wim *= K_ROULETTE;
                    sh.pp[g.rfoff] *= K_ROULETTE;
                    sh.pp[g.rfoff + 1] *= K_ROULETTE;
                }
            } else {
                launchflag = mediaidold & K_DET_MASK;
                relaunch = true;
                continue;
            }
        }

        if (REFLECT) {
            bool atface = true;

            if (FULL && g.svmc) {
                /* split voxels keep the properties of the side taken above; at the
                   plane, Fresnel reflection or refraction between the two labels
                   (src/mcx_core.cl:3072-3110) replaces the voxel-face test */
                if (hitintf) {
                    atface = false;

                    if (sh.prop[K_SV_LOWER(s.sv)].w != sh.prop[K_SV_UPPER(s.sv)].w) {
                        s.sn[0] = -s.sn[0];
                        s.sn[1] = -s.sn[1];
                        s.sn[2] = -s.sn[2];
                        s.sn[3] = -s.sn[3];

                        if (svmc_reflect(n1, v, s, sh.prop, rng)) {
                            launchflag = mediaidold & K_DET_MASK;
                            relaunch = true;
                            testint = true;
                            continue;
                        }
                    } else {
                        s.prop = sh.prop[K_SV_CURLABEL(s.sv)];
                    }
                }
            } else {
                s.prop = getprop<FULL>(g, sh, s.mediaid, s.idx1d);
            }

            if (atface && ((s.mediaid && g.doreflect) ||
                    (s.mediaid == 0 && (((isdet & 0xF) == 0 && g.doreflect) || ((isdet & 0xF) == kReflect ||
                                        (isdet & 0xF) == kMirror)))) &&
                    (((isdet & 0xF) == kMirror) || n1 != s.prop.w)) {
                float Rtotal = 1.f;
                float cphi, sphi, stheta, ctheta;
                float tmp0 = n1 * n1;
                float tmp1 = s.prop.w * s.prop.w;
                cphi = fabsf((face == 0) ? v.x : (face == 1 ? v.y : v.z));
                sphi = 1.f - cphi * cphi;
                f.z = 1.f - kdiv(tmp0, tmp1) * sphi;

                if (f.z > 0.f && (isdet & 0xF) != kMirror) {
                    ctheta = tmp0 * cphi * cphi + tmp1 * f.z;
                    stheta = 2.f * n1 * s.prop.w * cphi * ksqrt(f.z);
                    Rtotal = kdiv(ctheta - stheta, ctheta + stheta);
                    ctheta = tmp1 * cphi * cphi + tmp0 * f.z;
                    Rtotal = (Rtotal + kdiv(ctheta - stheta, ctheta + stheta)) * 0.5f;
                }

                if (Rtotal < 1.f && (!(s.mediaid == 0 && ((isdet & 0xF) == kMirror))) && rng.next() > Rtotal) {
                    transmit(v, n1, s.prop.w, face);

                    if (s.mediaid == 0) {
                        launchflag = (((s.idx1d == K_OUT_MAX && g.bc[9 + face]) ||
                                       (s.idx1d == K_OUT_MIN && g.bc[6 + face])) ? K_OUT_MIN : (mediaidold & K_DET_MASK));
                        relaunch = true;
                        continue;
                    }
                } else {
                    if (face == 0) {
                        v.x = -v.x;
                        p.x = nextafter_ref(rintf(p.x), 0);
                        s.fx = to_short((int)rintf(p.x));
                    } else if (face == 1) {
                        v.y = -v.y;
                        p.y = nextafter_ref(rintf(p.y), 0);
                        s.fy = to_short((int)rintf(p.y));
                    } else {
                        v.z = -v.z;
                        p.z = nextafter_ref(rintf(p.z), 0);
                        s.fz = to_short((int)rintf(p.z));
                    }

                    s.idx1d = idx1dold;
                    keepold = true;
                    /* back in the voxel it came from, whose label mediaidold holds
                       (the reference reloads it from the volume) */
                    s.mediaid = mediaidold & K_MED_MASK;

                    if (FULL && g.svmc) {
                        /* the side of the plane the reflected photon lands on; label 0
                           there ends it (src/mcx_core.cl:3221-3243) */
                        s.prop = svmc_prop(g, sh.prop, s.mediaid, s.idx1d, p, s.fx, s.fy, s.fz, s);

                        if (K_SV_CURLABEL(s.sv) == 0) {
                            launchflag = mediaidold & K_DET_MASK;
                            relaunch = true;
                            continue;
                        }
                    } else {
                        s.prop = getprop<FULL>(g, sh, s.mediaid, s.idx1d);
                    }

                    n1 = s.prop.w;
                    /* drop the boundary code so a later same-voxel step restores
                       the voxel's own label (the reference ORs the code in) */
                    isdet &= K_DET_MASK;
                }
            }

            if (FULL && g.moveonly && (s.mediaid == 0 || s.idx1d == K_OUT_MIN || s.idx1d == K_OUT_MAX)) {
                /* move-only trajectories (debug flag T, src/mcx_core.cl:3287-3291): the
                   lane's whole run ends here, without its energy write-back */
                moveend = true;
                break;
            }
        }
    }

    depcache_flush(g, sh, dcache);
    return !moveend;
}

#ifdef MCX_WAVES_PER_EU
#define MCX_KATTR __attribute__((amdgpu_waves_per_eu(MCX_WAVES_PER_EU)))
#else
#define MCX_KATTR
#endif
#if MCX_PARITY
#define MCX_KBOUND(SAVEDET, REFLECT, SIMPLE, FULL) MCX_MAX_BLOCK
#else
#define MCX_KBOUND(SAVEDET, REFLECT, SIMPLE, FULL) MCX_FAST_BOUND_FULL(SAVEDET, REFLECT, SIMPLE, FULL)
#endif
template <bool SAVEDET, bool REFLECT, bool SIMPLE, bool FULL>
__global__ void __launch_bounds__(MCX_KBOUND(SAVEDET, REFLECT, SIMPLE, FULL)) MCX_KATTR transport_kernel(const MCXKParam g) {
    extern __shared__ float4 lds[];
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t pplen = g.pplen;
    __shared__ uint32_t pool[2];
    Shared sh;
    sh.pool = pool;
    sh.prop = lds;
    sh.invcdf = (float*)(lds + g.medianum);
    sh.box = (BoxCell*)(sh.invcdf + ((g.nphase + 3) & ~3u));
    float* ppbase = (float*)(sh.box + g.box_cells);
    sh.repoff = (size_t)(idx & (int)(g.nrep - 1)) * g.fieldlen;

    for (uint32_t i = threadIdx.x; i < g.medianum; i += blockDim.x) {
        sh.prop[i] = g.prop[i];
    }

    for (uint32_t i = threadIdx.x; i < g.nphase; i += blockDim.x) {
        sh.invcdf[i] = g.invcdf[i];
    }

    for (uint32_t i = threadIdx.x; i < g.box_cells; i += blockDim.x) {
        sh.box[i] = (BoxCell)0;
    }

    if (g.dynamic && threadIdx.x == 0) {
        /* this workgroup's share: threadphoton photons, +1 for the first oddphoton
           groups.  The counter counts from 0 (run-time hand-out uses no photon
           ids), so the host only has to keep threadphoton + 1 + the lanes' one
           overshooting draw each below 2^32, whatever the total photon count. */
        pool[0] = 0u;
        pool[1] = g.threadphoton + (blockIdx.x < (uint32_t)g.oddphoton ? 1u : 0u);
    }

    /* lane-interleaved scratch: element k of lane t at [k*blockDim + t] would avoid
       bank conflicts, but the record copy reads it sequentially; a per-lane block
       keeps the indexing simple and the scratch is touched rarely */
    sh.pp = ppbase + threadIdx.x * pplen;

    for (uint32_t i = 0; i < pplen; i++) {
        sh.pp[i] = 0.f;
    }

    __syncthreads();

    /* lanes without photons skip the transport loop but still join the box flush */
    const bool active = g.dynamic ? (uint32_t)idx < g.nthread :
                        !((uint32_t)idx >= g.threadphoton * g.nthread + (uint32_t)g.oddphoton ||
                          (uint32_t)idx >= g.nthread);
    const uint32_t pd = g.partialdata;
    /* the tail workgroup may hold lanes past nthread: they read nothing */
    const uint32_t sidx = (uint32_t)idx < g.nthread ? (uint32_t)idx : 0u;
    sh.pp[pd] = g.genergy[sidx * 2];
    sh.pp[pd + 1] = g.genergy[sidx * 2 + 1];
    Rng rng;
    rng.s0 = (uint64_t)g.nseed[sidx * 4] << 32 | g.nseed[sidx * 4 + 1];
    rng.s1 = (uint64_t)g.nseed[sidx * 4 + 2] << 32 | g.nseed[sidx * 4 + 3];

    if (active && transport_lane<SAVEDET, REFLECT, SIMPLE, FULL>(g, sh, rng, idx)) {
        g.genergy[idx * 2] = sh.pp[pd];
        g.genergy[idx * 2 + 1] = sh.pp[pd + 1];
    }

    if (g.box_cells) {
        /* flush the workgroup's private box into its replica of the global grid */
        __syncthreads();
        const size_t boff = (size_t)(blockIdx.x & (g.nrep - 1)) * g.fieldlen;

        for (uint32_t i = threadIdx.x; i < g.box_cells; i += blockDim.x) {
            double val = (double)sh.box[i];

            if (val != 0.0) {
                uint32_t x = i % g.box_nx, r = i / g.box_nx;
                uint32_t y = r % g.box_ny;
                r /= g.box_ny;
                uint32_t z = r % g.box_nz, gate = r / g.box_nz;
                uint32_t vox = (g.box_z0 + z) * g.dimxy + (g.box_y0 + y) * g.dimx + (g.box_x0 + x);
                gate = (gate < g.box_ng) ? gate : gate - g.box_ng + g.box_imgate;
                atomicAdd(g.field + boff + vox + gate * g.nvox, val);
            }
        }
    }
}

/* sum of a double over the 64 lanes of a wave (result in every lane) */
__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) {
        v += __shfl_xor(v, off, 64);
    }

    return v;
}

/* Batched transport: many simulations in one persistent launch.
   Each instance has a photon counter {next, end} in global memory; lanes draw
   its photons one at a time with an atomic add, so any number of workgroups
   can share an instance and none waits for another at item boundaries.  The
   instances are split into nqueue lists, one per XCD: workgroup b starts at
   list b % nqueue (the XCD it is dispatched to), works on the list's first
   instance with photons left, and moves on when that one is drained; once its
   own list is drained it helps the others.  So an XCD's L2 holds few volumes
   and grids at a time.  On each instance a workgroup stages the tables in LDS,
   runs its lanes until the counter is exhausted, reduces the launched/escaped
   weight per wave into the instance's energy tally and flushes its deposit
   box.  Lanes keep their RNG streams across instances and write them back at
   the end, so the next launch continues them.  With cost != NULL the
   workgroup's wall-clock time per instance is summed per instance (the
   scheduler's cost probe).

   state: 2*nqueue words {head, end} of the lists in qlist, then 2 words
   {next photon, end} per instance. */
template <bool REFLECT>
/* four waves per SIMD: at most 128 VGPRs (the general transport path alone
   would take 140 and three waves) */
__global__ void __launch_bounds__(MCX_BATCH_BLOCK_MAX) __attribute__((amdgpu_waves_per_eu(4))) batch_kernel(const MCXKParam* __restrict__ prm,
        const uint32_t* __restrict__ qlist, const uint32_t nqueue, uint32_t* __restrict__ state,
        uint32_t* __restrict__ nseed, double* __restrict__ energy, unsigned long long* __restrict__ cost) {
    extern __shared__ float4 lds[];
    __shared__ uint32_t cur;
    uint32_t* const pools = state + 2 * nqueue;
    const int lane = blockIdx.x * blockDim.x + threadIdx.x;
    Rng rng;
    rng.s0 = (uint64_t)nseed[lane * 4] << 32 | nseed[lane * 4 + 1];
    rng.s1 = (uint64_t)nseed[lane * 4 + 2] << 32 | nseed[lane * 4 + 3];
    const uint32_t home = blockIdx.x % nqueue;

    for (;;) {
        if (threadIdx.x == 0) {
            uint32_t found = 0xFFFFFFFFu;

            /* every pass either finds an instance with photons left or advances a
               list head, so the search ends */
            for (uint32_t k = 0; k < nqueue && found == 0xFFFFFFFFu; k++) {
                uint32_t q = home + k < nqueue ? home + k : home + k - nqueue;

                for (;;) {
                    uint32_t h = __hip_atomic_load(state + 2 * q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);

                    if (h >= state[2 * q + 1]) {
                        break;
                    }

                    uint32_t inst = qlist[h];

                    if (__hip_atomic_load(pools + 2 * inst, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
                            pools[2 * inst + 1]) {
                        found = inst;
                        break;
                    }

                    atomicCAS(state + 2 * q, h, h + 1);
                }
            }

            cur = found;
        }

        __syncthreads();
        const uint32_t inst = __builtin_amdgcn_readfirstlane(cur);

        if (inst == 0xFFFFFFFFu) {
            break;
        }

        const MCXKParam& g = prm[inst];
        const long long tstart = cost ? wall_clock64() : 0;
        const uint32_t pplen = g.pplen;
        Shared sh;
        sh.pool = pools + 2 * inst;
        sh.prop = lds;
        sh.invcdf = (float*)(lds + g.medianum);
        sh.box = (BoxCell*)(sh.invcdf + ((g.nphase + 3) & ~3u));
        sh.pp = (float*)(sh.box + g.box_cells) + threadIdx.x * pplen;
        sh.repoff = 0;

        for (uint32_t i = threadIdx.x; i < g.medianum; i += blockDim.x) {
            sh.prop[i] = g.prop[i];
        }

        for (uint32_t i = threadIdx.x; i < g.nphase; i += blockDim.x) {
            sh.invcdf[i] = g.invcdf[i];
        }

        for (uint32_t i = threadIdx.x; i < g.box_cells; i += blockDim.x) {
            sh.box[i] = (BoxCell)0;
        }

        for (uint32_t i = 0; i < pplen; i++) {
            sh.pp[i] = 0.f;
        }

        __syncthreads();

        /* the general path only: with the pencil-beam special path inlined beside
           it the kernel needs more registers and runs fewer waves per SIMD */
        transport_lane<false, REFLECT, false, false>(g, sh, rng, lane);

        const uint32_t pd = g.partialdata;
        double esc = wave_sum((double)sh.pp[pd]);
        double tot = wave_sum((double)sh.pp[pd + 1]);

        if ((threadIdx.x & 63) == 0 && tot != 0.0) {
            atomicAdd(energy + 2 * inst, esc);
            atomicAdd(energy + 2 * inst + 1, tot);
        }

        __syncthreads();

        for (uint32_t i = threadIdx.x; i < g.box_cells; i += blockDim.x) {
            double val = (double)sh.box[i];

            if (val != 0.0) {
                uint32_t x = i % g.box_nx, r = i / g.box_nx;
                uint32_t y = r % g.box_ny;
                r /= g.box_ny;
                uint32_t z = r % g.box_nz, gate = r / g.box_nz;
                uint32_t vox = (g.box_z0 + z) * g.dimxy + (g.box_y0 + y) * g.dimx + (g.box_x0 + x);
                atomicAdd(g.field + vox + gate * g.nvox, val);
            }
        }

        __syncthreads();

        if (cost && threadIdx.x == 0) {
            atomicAdd(cost + inst, (unsigned long long)(wall_clock64() - tstart));
        }
    }

    nseed[lane * 4] = (uint32_t)(rng.s0 >> 32);
    nseed[lane * 4 + 1] = (uint32_t)rng.s0;
    nseed[lane * 4 + 2] = (uint32_t)(rng.s1 >> 32);
    nseed[lane * 4 + 3] = (uint32_t)rng.s1;
}

}  // namespace MCXK_NS

namespace mcxk = MCXK_NS;

"""ctypes wrapper of the CPU oracle (oracle/build/libmcxoracle.so), test use only.

The oracle is the checker: it restates the reference kernel on the CPU and is
compared against the gfx950 kernel.  It consumes a Config validated by the
product's configuration layer (mcxcl_amd.pmcxcl.prepare) and fills the same
outputs mcx_run_simulation does.
"""

import ctypes
import os

import numpy as np

from mcxcl_amd import pmcxcl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "build", "libmcxoracle.so")
# the CPU-baseline build of the same source (libm math, no counters, float32
# tallies, -O3): timed beside the GPU, never used as the checker
ORACLE_FAST_SO = os.path.join(ROOT, "oracle", "build", "libmcxoracle_fast.so")


class OracleCounters(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("nstep", "nscat", "ndep", "nbnd", "nlaunch", "nrand")]


class OracleResult(ctypes.Structure):
    _fields_ = [
        ("cnt", OracleCounters),
        ("rawfield", ctypes.c_void_p),
        ("fieldlen", ctypes.c_size_t),
        ("genergy", ctypes.c_void_p),
        ("nthread", ctypes.c_uint),
        ("seconds", ctypes.c_double),
    ]


def load(path=ORACLE_SO):
    if not os.path.exists(path):
        raise ImportError("oracle not built: run `make -C oracle`")
    lib = ctypes.CDLL(path)
    lib.oracle_run.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_int,
                               ctypes.POINTER(OracleResult)]
    lib.oracle_run.restype = ctypes.c_int
    lib.oracle_free_result.argtypes = [ctypes.POINTER(OracleResult)]
    lib.oracle_glibc_rand.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.c_void_p]
    lib.oracle_rng_stream.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_void_p, ctypes.c_void_p]
    lib.oracle_hitgrid.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    lib.oracle_hitgrid.restype = ctypes.c_float
    lib.oracle_reflectcoeff.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_float, ctypes.c_int]
    lib.oracle_reflectcoeff.restype = ctypes.c_float
    lib.oracle_rotatevector.argtypes = [ctypes.c_void_p] + [ctypes.c_float] * 4
    lib.oracle_fpmath.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    lib.oracle_result_sizeof.restype = ctypes.c_size_t
    lib.oracle_launch_samples.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_void_p]
    lib.oracle_launch_samples.restype = ctypes.c_int
    lib.oracle_pol_scatter_samples.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_void_p, ctypes.c_uint,
                                               ctypes.c_void_p]
    lib.oracle_scatter_samples.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                                           ctypes.c_void_p]
    assert lib.oracle_result_sizeof() == ctypes.sizeof(OracleResult)
    return lib


_LIB = None
_FAST = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = load()
    return _LIB


def fastlib():
    global _FAST
    if _FAST is None:
        _FAST = load(ORACLE_FAST_SO)
    return _FAST


def run(user, nthread, threadoffset=0, threadcount=0, nomp=0, fast=False):
    """Run the oracle on a pmcxcl-style config with the given persistent-thread count.

    Returns the pmcxcl output dict plus ``flux64`` (unnormalised double field),
    ``genergy`` (per-thread float tallies), ``counters`` and ``seconds``.
    ``fast`` runs the CPU-baseline build instead of the checker (no counters)."""
    L = fastlib() if fast else lib()
    with pmcxcl.prepare(user) as pc:
        res = OracleResult()
        rc = L.oracle_run(ctypes.addressof(pc.c), nthread, threadoffset, threadcount, nomp, ctypes.byref(res))
        if rc:
            raise RuntimeError("oracle_run failed (%d)" % rc)
        try:
            out = pc.collect()
            raw = np.ctypeslib.as_array((ctypes.c_double * res.fieldlen).from_address(res.rawfield)).copy()
            dims = [pc.c.dim.x, pc.c.dim.y, pc.c.dim.z, -1]
            out["flux64"] = raw.reshape(dims, order="F")
            out["genergy"] = np.ctypeslib.as_array(
                (ctypes.c_float * (2 * res.nthread)).from_address(res.genergy)).copy()
            out["counters"] = {k: getattr(res.cnt, k) for k, _ in OracleCounters._fields_}
            out["seconds"] = res.seconds
            out.update(adjoint_outputs(pc.c))
        finally:
            L.oracle_free_result(ctypes.byref(res))
    return out


def _cw_sum(field, slot, maxgate):
    """sum over the gates of one source slot, in gate order (float32)"""
    acc = np.zeros(field.shape[0], np.float32)
    for t in range(maxgate):
        acc = acc + field[:, slot * maxgate + t]
    return acc


def _fd_grad(c3, axis):
    """second-order differences along one axis in voxel units: central inside,
    one-sided at the faces, first order on a 2-voxel axis, 0 on a 1-voxel axis
    (src/mcx_core.cl:3344-3382), float32 in the reference's operation order"""
    n = c3.shape[axis]
    g = np.zeros_like(c3)
    if n <= 1:
        return g
    a = np.moveaxis(c3, axis, 0)
    r = np.moveaxis(g, axis, 0)
    h = np.float32(0.5)
    if n == 2:
        r[0] = a[1] - a[0]
        r[1] = a[1] - a[0]
        return g
    r[0] = (np.float32(-3.0) * a[0] + np.float32(4.0) * a[1] - a[2]) * h
    r[-1] = (a[-3] - np.float32(4.0) * a[-2] + np.float32(3.0) * a[-1]) * h
    r[1:-1] = (a[2:] - a[:-2]) * h
    return g


def adjoint_outputs(c):
    """The adjoint outputs from the oracle's normalised per-source field
    (src/mcx_core.cl:3399-3512 and their scaling, src/mcx_host.cpp:1467-1642),
    restated in numpy float32: {key: [Nx, Ny, Nz, Ns*Nd]} with pair s*Nd + d."""
    ot = c.outputtype[0]
    if ot < pmcxcl.OT_ADJOINT or not c.exportfield or not c.detnum or c.extrasrclen + 1 <= c.detnum:
        return {}
    dims = (c.dim.x, c.dim.y, c.dim.z)
    nvox = int(np.prod(dims))
    maxgate = int((c.tend - c.tstart) / c.tstep + 0.5)
    ns, nd = c.extrasrclen + 1 - c.detnum, c.detnum
    # RF forward: the imaginary half of the field follows the real one, and every
    # output is complex: re = re.re - im.im, im = re.im + im.re (src/mcx_core.cl:3430-3511)
    rf = c.omega > 0 and c.seed != pmcxcl.SEED_FROM_FILE
    half = nvox * maxgate * (ns + nd)
    flat = np.ctypeslib.as_array((ctypes.c_float * (half * (2 if rf else 1))).from_address(c.exportfield))
    parts = [flat[:half].reshape((nvox, -1), order="F")]
    if rf:
        parts.append(flat[half:].reshape((nvox, -1), order="F"))
    cws = [[_cw_sum(f, k, maxgate) for k in range(ns + nd)] for f in parts]
    gradss = [[[_fd_grad(v.reshape(dims, order="F"), ax).reshape(-1, order="F") for ax in range(3)] for v in cw]
              for cw in cws]
    vvox = np.float32(c.steps.x) * np.float32(c.steps.y) * np.float32(c.steps.z)
    prop = np.ctypeslib.as_array((ctypes.c_float * (4 * c.medianum)).from_address(c.prop)).reshape(-1, 4)
    labels = np.ctypeslib.as_array((ctypes.c_uint32 * nvox).from_address(c.vol)) & 0xFF

    def opscale(musp):
        op = np.zeros(256, np.float32)
        for m in range(c.medianum):
            mus, onemg = np.float32(prop[m, 1]), np.float32(1.0) - np.float32(prop[m, 2])
            if mus > 0 and onemg > 0:
                den = np.float32(3.0) * onemg * onemg * mus * mus if musp else np.float32(3.0) * onemg * mus * mus
                op[m] = np.float32(1.0) / den
        return op[labels]

    def block(kind, scale, op=None):
        out = np.empty((nvox, ns * nd), np.complex64 if rf else np.float32)
        for s in range(ns):
            for d in range(nd):
                if kind == 0:
                    cw = cws[0]
                    v = cw[s] * cw[ns + d]
                    if rf:
                        cwi = cws[1]
                        vi = cw[s] * cwi[ns + d] + cwi[s] * cw[ns + d]
                        v = v - cwi[s] * cwi[ns + d]
                else:
                    gs, gd = gradss[0][s], gradss[0][ns + d]
                    v = gs[0] * gd[0] + gs[1] * gd[1] + gs[2] * gd[2]
                    if rf:
                        gsi, gdi = gradss[1][s], gradss[1][ns + d]
                        vi = gs[0] * gdi[0] + gs[1] * gdi[1] + gs[2] * gdi[2] + gsi[0] * gd[0] + gsi[1] * gd[1] + \
                            gsi[2] * gd[2]
                        v = v - (gsi[0] * gdi[0] + gsi[1] * gdi[1] + gsi[2] * gdi[2])
                v = v * np.float32(scale)
                if op is not None:
                    v = v * op
                if rf:
                    vi = vi * np.float32(scale)
                    if op is not None:
                        vi = vi * op
                    v = v + 1j * vi
                out[:, s * nd + d] = v
        return out.reshape(dims + (ns * nd,), order="F")

    mua = lambda: block(0, -vvox)
    grad = lambda op=None: block(1, -np.float32(c.unitinmm), op)
    if ot == 11:
        return {"jmua": mua()}
    if ot == 12:
        return {"jd": grad()}
    if ot == 13:
        return {"jmus": grad(opscale(False))}
    if ot == 14:
        return {"jmusp": grad(opscale(True))}
    if ot == 15:
        return {"jmua": mua(), "jd": grad()}
    return {"jmua": mua(), "jmusp": grad(opscale(True))}


def glibc_rand(seed, n):
    out = np.zeros(n, dtype=np.uint32)
    lib().oracle_glibc_rand(seed, n, out.ctypes.data)
    return out


def rng_stream(seed4, n):
    s = np.asarray(seed4, dtype=np.uint32)
    out = np.zeros(n, dtype=np.float32)
    state = np.zeros(2, dtype=np.uint64)
    lib().oracle_rng_stream(s.ctypes.data, n, out.ctypes.data, state.ctypes.data)
    return out, state


def fpmath(which, x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros_like(x)
    lib().oracle_fpmath(which, x.ctypes.data, y.ctypes.data, x.size)
    return y


def hitgrid(p, v, ijk):
    p = np.asarray(p, dtype=np.float32)
    v = np.asarray(v, dtype=np.float32)
    ijk = np.asarray(ijk, dtype=np.int16)
    face = ctypes.c_int(0)
    d = lib().oracle_hitgrid(p.ctypes.data, v.ctypes.data, ijk.ctypes.data, ctypes.byref(face))
    return d, face.value


def reflectcoeff(v, n1, n2, flipdir):
    v = np.asarray(v, dtype=np.float32)
    return lib().oracle_reflectcoeff(v.ctypes.data, n1, n2, flipdir)


def rotatevector(v, stheta, ctheta, sphi, cphi):
    v = np.array(v, dtype=np.float32)
    lib().oracle_rotatevector(v.ctypes.data, stheta, ctheta, sphi, cphi)
    return v


def launch_samples(user, n):
    """n successive photon launches of thread 0 on a pmcxcl-style config:
    {"p": [n, 3], "w": [n], "v": [n, 3], "label": [n]} right after launch"""
    out = np.zeros((n, 8), np.float32)
    with pmcxcl.prepare(user) as pc:
        k = lib().oracle_launch_samples(ctypes.addressof(pc.c), n, out.ctypes.data)
    out = out[:k]
    return {"p": out[:, :3], "w": out[:, 3], "v": out[:, 4:7], "label": out[:, 7].astype(np.int64)}


def scatter_samples(user, label, nscat, n):
    """n scattering events in medium `label` after `nscat` earlier events:
    {"ctheta", "stheta", "cosdev"} (cosdev: cosine between old and new direction)"""
    out = np.zeros((n, 3), np.float32)
    with pmcxcl.prepare(user) as pc:
        lib().oracle_scatter_samples(ctypes.addressof(pc.c), label, nscat, n, out.ctypes.data)
    return {"ctheta": out[:, 0], "stheta": out[:, 1], "cosdev": out[:, 2]}


def pol_scatter_samples(user, label, stokes, n):
    """n scattering events in polarised medium `label` from direction (0.6, 0, 0.8)
    with the given Stokes vector: {"theta", "phi", "v" [n, 3], "stokes" [n, 4]}"""
    out = np.zeros((n, 9), np.float32)
    s4 = np.asarray(stokes, np.float32)
    with pmcxcl.prepare(user) as pc:
        lib().oracle_pol_scatter_samples(ctypes.addressof(pc.c), label, s4.ctypes.data, n, out.ctypes.data)
        table = np.ctypeslib.as_array((ctypes.c_float * (pc.c.polmedianum * 181 * 4)).from_address(pc.c.smatrix))
        table = table.copy().reshape(pc.c.polmedianum, 181, 4)
        prop = np.ctypeslib.as_array((ctypes.c_float * (4 * pc.c.medianum)).from_address(pc.c.prop)).copy()
    return {"theta": out[:, 0], "phi": out[:, 1], "v": out[:, 2:5], "stokes": out[:, 5:9], "smatrix": table,
            "prop": prop.reshape(-1, 4)}

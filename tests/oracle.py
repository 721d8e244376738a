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


def load():
    if not os.path.exists(ORACLE_SO):
        raise ImportError("oracle not built: run `make -C oracle`")
    lib = ctypes.CDLL(ORACLE_SO)
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
    assert lib.oracle_result_sizeof() == ctypes.sizeof(OracleResult)
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = load()
    return _LIB


def run(user, nthread, threadoffset=0, threadcount=0, nomp=0):
    """Run the oracle on a pmcxcl-style config with the given persistent-thread count.

    Returns the pmcxcl output dict plus ``flux64`` (unnormalised double field),
    ``genergy`` (per-thread float tallies), ``counters`` and ``seconds``."""
    L = lib()
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
        finally:
            L.oracle_free_result(ctypes.byref(res))
    return out


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

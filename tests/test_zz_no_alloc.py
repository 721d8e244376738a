"""mcx_run_simulation_noalloc: the forward run with the volume and outputs in
caller-owned device memory and every per-run buffer in the library's static
per-device scratch.  Checked through the C ABI against mcx_run_simulation on the
same validated Config: bit-reproducible mode to the parity bar (energies equal,
the float64 field within 1e-9 relative, the float output within 1e-6), on two
calls in a row, the second on a smaller volume with fewer photons; fast mode
launches every photon and agrees with mcx_run_simulation's fast run within
Monte Carlo error; runs beyond the scratch are refused with an error.

The caller's device buffers are allocated here with hipMalloc from the same
HIP runtime the library uses (the caller may allocate; the entry point does not)."""

import ctypes

import numpy as np
import pytest

import mcxcl_amd
from mcxcl_amd import _native, pmcxcl, workloads

pytestmark = pytest.mark.gpu

_hip = ctypes.CDLL("libamdhip64.so.7")
_hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
_hip.hipFree.argtypes = [ctypes.c_void_p]
_hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
_hip.hipDeviceSynchronize.argtypes = []
H2D, D2H = 1, 2

lib = _native.lib
lib.mcx_run_simulation_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
lib.mcx_run_simulation_noalloc.restype = ctypes.c_int


class DevBuf:
    def __init__(self, nbytes):
        self.ptr = ctypes.c_void_p()
        assert _hip.hipMalloc(ctypes.byref(self.ptr), max(nbytes, 1)) == 0
        self.nbytes = nbytes

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes == self.nbytes
        assert _hip.hipMemcpy(self.ptr, arr.ctypes.data, arr.nbytes, H2D) == 0

    def download(self, dtype, n):
        out = np.empty(n, dtype)
        assert _hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        _hip.hipFree(self.ptr)


def run_noalloc(cfg, with_fluence=True):
    """validate cfg as mcx_run_simulation would, stage its volume in device
    memory, run the no-allocation entry point, return (stat, field64, fluence)"""
    with pmcxcl.prepare(cfg) as pc:
        c = pc.c
        nvox, fieldlen = pc.nvox, pc.fieldlen
        media = np.ctypeslib.as_array((ctypes.c_uint32 * nvox).from_address(c.vol)).copy()
        dmedia, dfield = DevBuf(media.nbytes), DevBuf(fieldlen * 8)
        dflu = DevBuf(fieldlen * 4) if with_fluence else None
        try:
            dmedia.upload(media)
            dfield.upload(np.full(fieldlen, np.nan))      # the call clears it
            rc = lib.mcx_run_simulation_noalloc(ctypes.addressof(c), dmedia.ptr, nvox, dfield.ptr,
                                                dflu.ptr if dflu else None, fieldlen, None)
            if rc:
                raise RuntimeError(_native.last_error())
            stat = {k: getattr(c, k) for k in ("energytot", "energyesc", "energyabs", "normalizer")}
            f64 = dfield.download(np.float64, fieldlen)
            flu = dflu.download(np.float32, fieldlen) if dflu else None
        finally:
            for b in (dmedia, dfield, dflu):
                if b is not None:
                    b.free()
    return stat, f64, flu


def _compare_parity(cfg):
    stat, f64, flu = run_noalloc(cfg)
    ref = mcxcl_amd.run(dict(cfg, rawfield=1))
    assert stat["energytot"] == ref["stat"]["energytot"] == float(cfg["nphoton"])
    assert stat["energyabs"] == ref["stat"]["energyabs"]
    assert stat["normalizer"] == ref["stat"]["normalizer"]
    np.testing.assert_allclose(f64, ref["flux64"].ravel(order="F"), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(flu, ref["flux"].ravel(order="F"), rtol=1e-6, atol=0)
    assert np.isfinite(f64).all() and f64.sum() > 0


def test_noalloc_matches_run_simulation_twice_then_smaller():
    big = workloads.sphere_cube_workload(nphoton=200000, issavedet=0, autopilot=0, nthread=4096, nblocksize=64,
                                         mathmode=1)
    small = workloads.homogeneous_cube_workload(nphoton=50000, vol=np.ones((40, 40, 40), np.uint8),
                                                srcpos=[19, 19, 0], autopilot=0, nthread=2048, nblocksize=64,
                                                mathmode=1)
    _compare_parity(big)
    _compare_parity(small)


def test_noalloc_fast_mode_launches_every_photon():
    cfg = workloads.homogeneous_cube_workload(nphoton=2000000, mathmode=0)
    stat, f64, flu = run_noalloc(cfg)
    ref = mcxcl_amd.run(dict(cfg, rawfield=1))
    assert stat["energytot"] == ref["stat"]["energytot"] == 2e6
    a, b = stat["energyabs"] / stat["energytot"], ref["stat"]["energyabs"] / ref["stat"]["energytot"]
    assert abs(a - b) < 2e-3
    assert abs(f64.sum() / ref["flux64"].sum() - 1) < 5e-3
    np.testing.assert_allclose(flu.sum(), ref["flux"].sum(), rtol=5e-3)


def test_noalloc_refuses_what_its_scratch_cannot_hold():
    # bit-reproducible autopilot: 2560 lanes per CU, more than the scratch holds
    cfg = workloads.homogeneous_cube_workload(nphoton=1000, mathmode=1)
    with pytest.raises(RuntimeError, match="lanes"):
        run_noalloc(cfg)
    with pytest.raises(RuntimeError, match="detected photons"):
        run_noalloc(workloads.sphere_cube_workload(nphoton=1000, autopilot=0, nthread=1024, nblocksize=64))

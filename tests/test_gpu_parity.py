"""GPU parity: the gfx950 kernel against the CPU oracle on the same seeds and thread map.

Bar: bit-exact for the RNG, the parity-mode arithmetic, detected-photon records,
seeds and the per-thread energy tallies; the double-precision fluence accumulator
agrees to 1e-9 relative (atomic summation order only) and the float output to
1 ulp-scale (rtol 1e-6).  All runs go through the C ABI (mcx_run_simulation).
"""

import numpy as np
import pytest

import mcxcl_amd
from mcxcl_amd import _native, workloads
import oracle
from helpers import sorted_records, sorted_seeds

pytestmark = pytest.mark.gpu

lib = _native.lib


def _device_fpmath(which, mode, x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    n = x.size // 2 if which == 6 else x.size
    y = np.zeros(n, dtype=np.float32)
    rc = lib.mcx_device_fpmath(0, which, mode, x.ctypes.data, y.ctypes.data, n)
    assert rc == 0, "device fpmath failed (%d)" % rc
    return y


@pytest.mark.parametrize("which,lo,hi", [(0, -100.0, 80.0), (1, 1e-30, 1e4), (2, 0.0, 6.3), (3, 0.0, 6.3),
                                         (4, -1.0, 1.0), (5, 1e-6, 1e6)])
def test_parity_math_bitexact(which, lo, hi):
    rng = np.random.default_rng(which)
    x = rng.uniform(lo, hi, 1 << 20).astype(np.float32)
    if which == 1:
        x = np.exp(rng.uniform(np.log(lo), np.log(hi), x.size)).astype(np.float32)
    gpu = _device_fpmath(which, 1, x)
    cpu = oracle.fpmath(which, x)
    assert np.array_equal(gpu.view(np.uint32), cpu.view(np.uint32))


def test_ieee_divide_sqrt_bitexact():
    rng = np.random.default_rng(7)
    a = rng.uniform(-1e3, 1e3, 1 << 20).astype(np.float32)
    b = rng.uniform(1e-3, 1e3, 1 << 20).astype(np.float32)
    q = _device_fpmath(6, 1, np.concatenate([a, b]))
    assert np.array_equal(q.view(np.uint32), (a / b).view(np.uint32))
    s = _device_fpmath(7, 1, np.abs(a))
    assert np.array_equal(s.view(np.uint32), np.sqrt(np.abs(a)).view(np.uint32))


def test_fast_math_accuracy():
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 6.28, 1 << 16).astype(np.float32)
    assert np.max(np.abs(_device_fpmath(2, 0, x) - np.sin(x.astype(np.float64)))) < 2e-6
    x = rng.uniform(-20, 0, 1 << 16).astype(np.float32)
    np.testing.assert_allclose(_device_fpmath(0, 0, x), np.exp(x.astype(np.float64)), rtol=4e-6)
    x = rng.uniform(1e-7, 1, 1 << 16).astype(np.float32)
    np.testing.assert_allclose(_device_fpmath(1, 0, x), np.log(x.astype(np.float64)), rtol=4e-6, atol=1e-6)


def test_rng_stream_bitexact():
    seed4 = np.array([0x12345678, 0x9ABCDEF0, 0x0F1E2D3C, 0x4B5A6978], dtype=np.uint32)
    out = np.zeros(4096, dtype=np.float32)
    assert lib.mcx_device_rng_stream(0, seed4.ctypes.data, out.size, out.ctypes.data) == 0
    ref, _ = oracle.rng_stream(seed4, out.size)
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))


def test_rng_dump_mode_matches_oracle():
    """debug flag R: field[i] = stream of thread i % nthread (src/mcx_core.cl:2408-2414)"""
    cfg = workloads.homogeneous_cube_workload(nphoton=1000, autopilot=0, nthread=4096, nblocksize=64,
                                              debuglevel="R", mathmode=1)
    gpu = mcxcl_amd.run(cfg)["flux"]
    ref = oracle.run(cfg, nthread=4096)["flux"]
    assert np.array_equal(gpu.view(np.uint32), ref.view(np.uint32))


def _parity_pair(cfg, nthread):
    cfg = dict(cfg, autopilot=0, nthread=nthread, nblocksize=64, mathmode=1, rawfield=1)
    gpu = mcxcl_amd.run(cfg)
    ref = oracle.run(cfg, nthread=nthread)
    return gpu, ref


@pytest.mark.parametrize("nphoton,nthread", [(20000, 2048), (100000, 6400)])
def test_homogeneous_cube_parity(nphoton, nthread):
    gpu, ref = _parity_pair(workloads.homogeneous_cube_workload(nphoton=nphoton), nthread)
    assert gpu["stat"]["energytot"] == ref["stat"]["energytot"]
    assert gpu["stat"]["energyabs"] == ref["stat"]["energyabs"]
    np.testing.assert_allclose(gpu["flux64"], ref["flux64"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(gpu["flux"], ref["flux"], rtol=1e-6, atol=0)


def test_sphere_detectors_parity():
    cfg = workloads.sphere_cube_workload(nphoton=100000, issaveseed=1, savedetflag="DSPMXVW")
    gpu, ref = _parity_pair(cfg, 4096)
    assert gpu["stat"]["energytot"] == ref["stat"]["energytot"]
    assert gpu["stat"]["energyabs"] == ref["stat"]["energyabs"]
    np.testing.assert_allclose(gpu["flux64"], ref["flux64"], rtol=1e-9, atol=1e-12)
    assert gpu["detp"].shape == ref["detp"].shape and gpu["detp"].shape[1] > 100
    assert np.array_equal(sorted_records(gpu["detp"]), sorted_records(ref["detp"]))
    assert np.array_equal(sorted_seeds(gpu["seeds"]), sorted_seeds(ref["seeds"]))


@pytest.mark.parametrize("srctype,extra", [
    ("isotropic", {"srcpos": [29.5, 29.5, 29.5]}),
    ("disk", {"srcparam1": [5, 0, 0, 0]}),
    ("planar", {"srcpos": [10, 10, 0], "srcparam1": [40, 0, 0, 0], "srcparam2": [0, 40, 0, 0]}),
    ("cone", {"srcparam1": [0.5, 0, 0, 0]}),
    ("gaussian", {"srcparam1": [3, 0, 0, 0]}),
])
def test_source_types_parity(srctype, extra):
    cfg = workloads.homogeneous_cube_workload(nphoton=20000, srctype=srctype, **extra)
    gpu, ref = _parity_pair(cfg, 2048)
    assert gpu["stat"]["energytot"] == ref["stat"]["energytot"]
    assert gpu["stat"]["energyabs"] == ref["stat"]["energyabs"]
    np.testing.assert_allclose(gpu["flux64"], ref["flux64"], rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("name,extra", [
    ("focused pencil", {"srcdir": [0, 0, 1, 20.0]}),
    ("detector flags on the z=0 face", {"bc": "______001000", "issavedet": 1, "detpos": [[29, 29, 0, 3]]}),
    ("pencil from outside the grid", {"srcpos": [29.5, 29.5, -3.0]}),
    ("diffuse reflectance into an air layer", "airlayer"),
])
def test_general_pencil_paths_parity(name, extra):
    """pencil beams that take the general launch/deposit code (not the specialised one)"""
    if extra == "airlayer":
        vol = np.ones((40, 40, 40), np.uint8)
        vol[:, :, :2] = 0
        cfg = dict(nphoton=20000, vol=vol, prop=[[0, 0, 1, 1], [0.005, 1.0, 0.01, 1.37]], srcpos=[20, 20, 0],
                   srcdir=[0, 0, 1], tstart=0, tend=5e-9, tstep=5e-9, issaveref=1, isreflect=1, seed=1648335518)
    else:
        cfg = workloads.homogeneous_cube_workload(nphoton=20000, **extra)
    gpu, ref = _parity_pair(cfg, 2048)
    assert gpu["stat"]["energytot"] == ref["stat"]["energytot"]
    assert gpu["stat"]["energyabs"] == ref["stat"]["energyabs"]
    np.testing.assert_allclose(gpu["flux64"], ref["flux64"], rtol=1e-9, atol=1e-12)
    if "detp" in ref:
        assert np.array_equal(sorted_records(gpu["detp"]), sorted_records(ref["detp"]))
    if extra == "airlayer":
        assert (ref["flux64"] < 0).sum() > 100   # reflectance deposits happened


@pytest.mark.parametrize("bc", ["aaaaaa", "rrrrrr", "mmmmmm", "ccrrrr", "______111111"])
def test_boundary_conditions_parity(bc):
    cfg = workloads.homogeneous_cube_workload(nphoton=10000, bc=bc, issavedet=1 if "1" in bc else 0)
    gpu, ref = _parity_pair(cfg, 2048)
    assert gpu["stat"]["energytot"] == ref["stat"]["energytot"]
    assert gpu["stat"]["energyabs"] == ref["stat"]["energyabs"]
    np.testing.assert_allclose(gpu["flux64"], ref["flux64"], rtol=1e-9, atol=1e-12)
    if "1" in bc:
        assert np.array_equal(sorted_records(gpu["detp"]), sorted_records(ref["detp"]))


def test_fast_mode_statistics():
    """hardware-transcendental build with run-time photon scheduling: every photon
    launched exactly once, same physics within Monte Carlo error"""
    cfg = workloads.homogeneous_cube_workload(nphoton=1000003)
    fast = mcxcl_amd.run(dict(cfg, mathmode=0))
    par = mcxcl_amd.run(dict(cfg, mathmode=1))
    assert fast["stat"]["energytot"] == 1000003.0
    assert par["stat"]["energytot"] == 1000003.0
    af = fast["stat"]["energyabs"] / fast["stat"]["energytot"]
    ap = par["stat"]["energyabs"] / par["stat"]["energytot"]
    assert abs(af - ap) < 3e-3, (af, ap)
    assert 0.26 < af < 0.285
    ff, fp = fast["flux"].sum(), par["flux"].sum()
    assert abs(ff - fp) / fp < 5e-3


@pytest.mark.parametrize("replaydet", [0, -1, 3])
def test_replay_jacobian_parity(replaydet):
    """replay of detected photons (saved seeds) with the Jacobian output:
    device equals oracle on the same inputs, and the float64 total equals
    sum_i w_i L_i of the recorded partial path lengths"""
    base = workloads.sphere_cube_workload(nphoton=100000, issaveseed=1, savedetflag="DP", tstep=5e-9)
    fwd = mcxcl_amd.run(dict(base, autopilot=0, nthread=4096, nblocksize=64, mathmode=1))
    detp, seeds = fwd["detp"], fwd["seeds"]
    assert detp.shape[1] > 200
    rep = dict(base, seed=seeds, detphotons=detp, outputtype="jacobian", replaydet=replaydet)
    gpu, ref = _parity_pair(rep, 4096)
    np.testing.assert_allclose(gpu["flux64"], ref["flux64"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(gpu["flux"], ref["flux"], rtol=1e-6, atol=1e-12)
    det = detp[0].astype(np.int64) & 0xFFFF
    L = detp[1:3].astype(np.float64)
    w = np.exp(-(np.array([0.005, 0.002])[:, None] * L).sum(0))
    sel = np.ones(det.size, bool) if replaydet <= 0 else det == replaydet
    np.testing.assert_allclose(gpu["flux64"].sum(), (w[sel] * L[:, sel].sum(0)).sum(), rtol=1e-6)


def test_replay_fast_mode_retraces_paths():
    """the hardware-math build replays its own detected photons along the same paths"""
    base = workloads.sphere_cube_workload(nphoton=1000000, issaveseed=1, savedetflag="DP", tstep=5e-9)
    fwd = mcxcl_amd.run(dict(base, mathmode=0))
    detp = fwd["detp"]
    assert detp.shape[1] > 1000
    rep = mcxcl_amd.run(dict(base, seed=fwd["seeds"], detphotons=detp, outputtype="jacobian", replaydet=0,
                             mathmode=0, rawfield=1))
    L = detp[1:3].astype(np.float64)
    w = np.exp(-(np.array([0.005, 0.002])[:, None] * L).sum(0))
    np.testing.assert_allclose(rep["flux64"].sum(), (w * L.sum(0)).sum(), rtol=1e-5)


def test_pervoxel_media_parity():
    """per-voxel mua/mus field: device equals oracle, exit weights included"""
    base = workloads.sphere_cube_workload(nphoton=50000, issaveseed=1, savedetflag="DPX", tstep=5e-10)
    vol = base["vol"]
    prop = np.asarray(base["prop"], np.float32)
    ii, jj, kk = np.meshgrid(np.arange(60), np.arange(60), np.arange(60), indexing="ij")
    pert = 1 + 0.3 * np.sin(ii / 7.0) * np.cos(jj / 9.0) * np.sin(kk / 5.0)
    mm = np.stack([prop[vol, 0] * pert, prop[vol, 1] * (2 - pert)]).astype(np.float32)
    gpu, ref = _parity_pair(dict(base, muamus=mm, savedetweight=1), 4096)
    assert gpu["stat"]["energytot"] == ref["stat"]["energytot"]
    assert gpu["stat"]["energyabs"] == ref["stat"]["energyabs"]
    np.testing.assert_allclose(gpu["flux64"], ref["flux64"], rtol=1e-9, atol=1e-12)
    a = np.vstack([gpu["detp"], gpu["detw"][None, :]])
    b = np.vstack([ref["detp"], ref["detw"][None, :]])
    assert np.array_equal(sorted_records(a), sorted_records(b))


def test_float32_raw_export_is_the_converted_accumulator():
    """Config.exportfield32: the float64 accumulator converted on the device
    (round to nearest) equals the float64 read-back rounded on the host, with
    and without the normalised float32 copy beside it"""
    base = workloads.sphere_cube_workload(nphoton=200000, issaveseed=1, savedetflag="DP", tstep=5e-9)
    fwd = mcxcl_amd.run(dict(base, mathmode=0))
    rep = dict(base, seed=fwd["seeds"], detphotons=fwd["detp"], outputtype="jacobian", replaydet=-1, mathmode=0)
    both = mcxcl_amd.run(dict(rep, rawfield=1, rawfield32=1))
    assert both["flux32"].dtype == np.float32 and both["flux32"].shape == both["flux64"].shape
    assert np.array_equal(both["flux32"], both["flux64"].astype(np.float32))
    assert "flux" in both
    only = mcxcl_amd.run(dict(rep, rawfield32=1, rawonly=1))
    assert "flux" not in only and "flux64" not in only
    assert only["flux32"].shape == both["flux32"].shape and only["flux32"].sum() > 0


def test_launch_shape_model():
    """the fast path's launch-shape model: the pencil-beam variant fits in 128
    VGPRs, so one 1024-lane workgroup (16 waves) fits a CU and four 256-lane
    ones do; with 150 KB of LDS only one workgroup fits"""
    import ctypes
    lib = _native.lib
    lib.mcx_fast_kernel_attr.argtypes = [ctypes.c_int] * 3 + [ctypes.POINTER(ctypes.c_int)] * 2
    lib.mcx_fast_occupancy.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    vgpr, lanes = ctypes.c_int(0), ctypes.c_int(0)
    mcxcl_amd.run(workloads.sphere_cube_workload(nphoton=10000))          # device and code object loaded
    assert lib.mcx_fast_kernel_attr(1, 1, 1, ctypes.byref(vgpr), ctypes.byref(lanes)) == 0
    print("pencil-beam variant: %d VGPRs, %d lanes at most" % (vgpr.value, lanes.value), flush=True)
    assert 32 <= vgpr.value <= 128 and lanes.value >= 1024
    assert lib.mcx_fast_occupancy(1024, 20000, 1, 1, 1) == 1
    assert lib.mcx_fast_occupancy(256, 20000, 1, 1, 1) == 4
    assert lib.mcx_fast_occupancy(1024, 150 * 1024, 1, 1, 1) == 1
    assert lib.mcx_fast_occupancy(256, 150 * 1024, 1, 1, 1) == 1

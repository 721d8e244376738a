"""More device parity: the kernel branches the core parity file does not
reach.  The remaining source types (fourier, arcsine, zgaussian, line, slit,
pencilarray, ring; src/mcx_core.cl:1632-2140), the other output types
(fluence, energy, pathlength; :2869-2905), unnormalised output, the
similarity rule (gscatter, :2484-2497), Russian roulette (:3031-3061), a 2D
domain (:2497-2509), a sampled phase function (invcdf, :2459-2475), a voxel
size other than 1 mm, uint16 label volumes, and replay WP and DCS deposits
with momentum transfer saved (:2568-2623).

Bar, as in test_gpu_parity.py: energies and records bit-exact, the float64
grid within 1e-9 relative (atomic summation order only), the float output
within 1e-6."""

import numpy as np
import pytest

import mcxcl_amd
import oracle
from helpers import sorted_records
from mcxcl_amd import workloads

pytestmark = pytest.mark.gpu

NT = 2048


def _pair(cfg, nthread=NT):
    cfg = dict(cfg, autopilot=0, nthread=nthread, nblocksize=64, mathmode=1, rawfield=1)
    return mcxcl_amd.run(cfg), oracle.run(cfg, nthread=nthread)


def _same(gpu, ref):
    assert gpu["stat"]["energytot"] == ref["stat"]["energytot"]
    assert gpu["stat"]["energyabs"] == ref["stat"]["energyabs"]
    assert gpu["flux"].shape == ref["flux"].shape
    np.testing.assert_allclose(gpu["flux64"], ref["flux64"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(gpu["flux"], ref["flux"], rtol=1e-6, atol=0)
    if "detp" in ref:
        assert np.array_equal(sorted_records(gpu["detp"]), sorted_records(ref["detp"]))


SOURCES = [
    ("fourier", {"srcpos": [10, 10, 0], "srcparam1": [40, 0, 0, 2.25], "srcparam2": [0, 40, 0, 1.5]}),
    ("arcsine", {"srcpos": [29.5, 29.5, 29.5]}),
    ("zgaussian", {"srcparam1": [0.3, 0, 0, 0]}),
    ("line", {"srcpos": [20, 29.5, 29.5], "srcparam1": [20, 0, 0, 0]}),
    ("line", {"srcpos": [20, 29.5, 29.5], "srcdir": [0, 1, 0], "srcparam1": [20, 0, 0, 0], "srcparam2": [0.5, 0, 0, 0]}),
    ("slit", {"srcpos": [20, 29.5, 0], "srcparam1": [20, 0, 0, 0]}),
    ("slit", {"srcpos": [20, 29.5, 0], "srcparam1": [20, 0, 0, 0], "srcparam2": [0.1, 0.2, 0, 0]}),
    ("pencilarray", {"srcpos": [10, 10, 0], "srcparam1": [40, 0, 0, 4], "srcparam2": [0, 40, 0, 3]}),
    ("ring", {"srcparam1": [8, 4, 0, 0]}),
    ("ring", {"srcparam1": [8, 4, 0.5, 2.5]}),
]


@pytest.mark.parametrize("srctype,extra", SOURCES, ids=["%s%d" % (s, i) for i, (s, _) in enumerate(SOURCES)])
def test_more_source_types(srctype, extra):
    cfg = workloads.homogeneous_cube_workload(nphoton=20000, srctype=srctype, **extra)
    gpu, ref = _pair(cfg)
    assert ref["stat"]["energytot"] > 0
    _same(gpu, ref)


OPTIONS = [
    ("fluence", {"outputtype": "fluence"}),
    ("energy", {"outputtype": "energy"}),
    ("length", {"outputtype": "length"}),
    ("unnormalised", {"isnormalized": 0}),
    ("gscatter", {"gscatter": 3, "prop": [[0, 0, 1, 1], [0.005, 1.0, 0.9, 1.37]]}),
    ("roulette", {"minenergy": 0.05}),
    ("voxel 0.5 mm", {"unitinmm": 0.5}),
    ("uint16 labels", {"vol": np.ones([60, 60, 60], dtype=np.uint16)}),
]


@pytest.mark.parametrize("name,extra", OPTIONS, ids=[n for n, _ in OPTIONS])
def test_output_types_and_options(name, extra):
    cfg = workloads.sphere_cube_workload(nphoton=40000, savedetflag="DSPX", **extra)
    if "prop" in extra:
        cfg["vol"] = np.ones([60, 60, 60], dtype=np.uint8)
    gpu, ref = _pair(cfg)
    _same(gpu, ref)


def test_two_dimensional_domain():
    cfg = workloads.homogeneous_cube_workload(nphoton=20000, vol=np.ones([60, 60, 1], dtype=np.uint8),
                                              srcpos=[29.5, 0, 0.5], srcdir=[0, 1, 0])
    gpu, ref = _pair(cfg)
    _same(gpu, ref)


def test_sampled_phase_function():
    """a tabulated inverse CDF of cos(theta) replaces Henyey-Greenstein"""
    invcdf = np.linspace(-1, 1, 1001)
    invcdf = np.sign(invcdf) * np.abs(invcdf) ** 0.7    # forward-peaked, monotone in [-1, 1]
    cfg = workloads.homogeneous_cube_workload(nphoton=20000, invcdf=invcdf[1:-1].astype(np.float32))
    gpu, ref = _pair(cfg)
    _same(gpu, ref)


@pytest.mark.parametrize("outputtype", ["wp", "wm"], ids=["wp", "dcs"])
def test_replay_scattering_deposits(outputtype):
    """'wp': replay weight at every scattering event; 'wm' (DCS): the same
    weighted by the momentum transfer 1 - cos(theta)"""
    base = workloads.sphere_cube_workload(nphoton=100000, issaveseed=1, savedetflag="DSPM", ismomentum=1,
                                          tstep=5e-9)
    gpu_f, ref_f = _pair(base)
    assert np.array_equal(sorted_records(gpu_f["detp"]), sorted_records(ref_f["detp"]))
    rep = dict(base, seed=ref_f["seeds"], detphotons=ref_f["detp"], outputtype=outputtype, replaydet=0)
    gpu, ref = _pair(rep)
    assert ref["flux64"].sum() > 0
    np.testing.assert_allclose(gpu["flux64"], ref["flux64"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(gpu["flux"], ref["flux"], rtol=1e-6, atol=0)


def test_fast_mode_with_many_media_and_full_records():
    """12 media with every record column: a 1024-lane workgroup's lane scratch
    would not fit in LDS, so the fast path runs 256-lane workgroups; it still
    agrees with the bit-reproducible run within Monte Carlo error"""
    vol = np.ones([60, 60, 60], dtype=np.uint8)
    for k in range(12):
        vol[:, :, 5 * k:5 * (k + 1)] = k + 1
    prop = [[0, 0, 1, 1]] + [[0.004 + 0.001 * k, 1.0, 0.8, 1.37] for k in range(12)]
    cfg = workloads.sphere_cube_workload(nphoton=200000, vol=vol, prop=prop, savedetflag="DSPMXVW", ismomentum=1,
                                         issaveseed=1, tstep=5e-9)
    fast = mcxcl_amd.run(dict(cfg, mathmode=0))
    slow = mcxcl_amd.run(dict(cfg, mathmode=1))
    n = cfg["nphoton"]
    assert fast["stat"]["energytot"] == n and slow["stat"]["energytot"] == n
    af, ap = (o["stat"]["energyabs"] / n for o in (fast, slow))
    assert abs(af - ap) <= 6 * np.sqrt(2 * ap * (1 - ap) / n), (af, ap)
    nf, ns = fast["detp"].shape[1], slow["detp"].shape[1]
    assert fast["detp"].shape[0] == slow["detp"].shape[0]
    assert abs(nf - ns) <= 6 * np.sqrt(2 * ns) + 5, (nf, ns)

"""Device parity of the full general path: the launches and outputs the common
general path leaves out, each against the oracle with the same seeds and
thread map (bit-reproducible mode).

Launches: pattern (one pattern scaling the weight, and two patterns sharing
each photon into interleaved slots), pattern3d, fourierx, fourierx2d, the
hyperboloid Gaussian beam and the angle-CDF launch in both its step and
interpolated forms (src/mcx_core.cl:1656-1852, :2098-2125).  Outputs: the RF
replay outputs r and s (real and imaginary halves) and the time-weighted
replay outputs t and b, with one slice per detector (replaydet -1) and pooled
(src/mcx_core.cl:2257-2263, :2568-2623, :2845-2858).

Bar, as in test_gpu_parity.py: energies and records bit-exact, the float64
grid within 1e-9 relative (atomic summation order only), the float output
within 1e-6.  One fast-mode run of the full path agrees with the
bit-reproducible one within Monte Carlo error."""

import math

import numpy as np
import pytest

import mcxcl_amd
import oracle
from helpers import sorted_records
from mcxcl_amd import workloads

pytestmark = pytest.mark.gpu

NT = 2048
RNG = np.random.default_rng(2024)
PAT2 = RNG.uniform(0.2, 1.0, (4, 3)).astype(np.float32)
PAT3 = RNG.uniform(0.2, 1.0, (4, 3, 2)).astype(np.float32)
PATS = RNG.uniform(0.2, 1.0, (2, 4, 3)).astype(np.float32)


def _pair(cfg, nthread=NT):
    cfg = dict(cfg, autopilot=0, nthread=nthread, nblocksize=64, mathmode=1, rawfield=1)
    return mcxcl_amd.run(cfg), oracle.run(cfg, nthread=nthread)


def _same(gpu, ref):
    assert gpu["stat"]["energytot"] == ref["stat"]["energytot"]
    assert gpu["stat"]["energyabs"] == ref["stat"]["energyabs"]
    assert gpu["stat"]["normalizer"] == ref["stat"]["normalizer"]
    assert gpu["flux"].shape == ref["flux"].shape
    np.testing.assert_allclose(gpu["flux64"], ref["flux64"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(gpu["flux"], ref["flux"], rtol=1e-6, atol=0)
    if "detp" in ref:
        assert np.array_equal(sorted_records(gpu["detp"]), sorted_records(ref["detp"]))


LAUNCHES = [
    ("pattern", {"srctype": "pattern", "srcpos": [10, 15, 0], "srcparam1": [40, 0, 0, 4],
                 "srcparam2": [0, 30, 0, 3], "srcpattern": PAT2}),
    ("pattern-shared", {"srctype": "pattern", "srcpos": [10, 15, 0], "srcparam1": [40, 0, 0, 4],
                        "srcparam2": [0, 30, 0, 3], "srcnum": 2, "srcpattern": PATS}),
    ("pattern3d", {"srctype": "pattern3d", "srcpos": [20, 20, 0], "srcparam1": [4, 3, 2, 0], "srcpattern": PAT3}),
    ("fourierx", {"srctype": "fourierx", "srcpos": [10, 10, 0], "srcparam1": [40, 0, 0, 20],
                  "srcparam2": [2, 1, 0.25, 0.5]}),
    ("fourierx2d", {"srctype": "fourierx2d", "srcpos": [10, 10, 0], "srcparam1": [40, 0, 0, 20],
                    "srcparam2": [2, 3, 0.1, 0.3]}),
    ("hyperboloid", {"srctype": "hyperboloid", "srcpos": [30, 30, 0], "srcparam1": [3, 10, 20, 0]}),
    ("angle-interp", {"srcpos": [29.5, 29.5, 0], "angleinvcdf": np.linspace(0, 0.45, 9).astype(np.float32)}),
    ("angle-step", {"srcpos": [29.5, 29.5, 0], "srcdir": [0, 0, 1, 1],
                    "angleinvcdf": np.linspace(0, 0.45, 9).astype(np.float32)}),
]


@pytest.mark.parametrize("name,extra", LAUNCHES, ids=[n for n, _ in LAUNCHES])
def test_full_path_launches(name, extra):
    cfg = workloads.sphere_cube_workload(nphoton=30000, savedetflag="DSPXVW", issaveseed=1, tstep=1e-9, **extra)
    gpu, ref = _pair(cfg)
    assert ref["stat"]["energytot"] > 0 and ref["flux64"].sum() > 0
    _same(gpu, ref)


def test_shared_patterns_with_reflectance():
    """issaveref = 1 with two patterns sharing photons: escaping weights go to the
    interleaved slot of each pattern (src/mcx_core.cl:1525-1540)"""
    vol = workloads.sphere_volume().copy()
    vol[:, :, 0] = 0                                       # a void layer over the tissue
    extra = dict(LAUNCHES)["pattern-shared"]
    cfg = workloads.sphere_cube_workload(nphoton=30000, savedetflag="DSPXVW", issaveseed=1, tstep=1e-9, vol=vol,
                                         issaveref=1, **extra)
    gpu, ref = _pair(cfg)
    assert (ref["flux64"] < 0).sum() > 100
    _same(gpu, ref)


@pytest.fixture(scope="module")
def forward():
    base = workloads.sphere_cube_workload(nphoton=100000, issaveseed=1, savedetflag="DSP", tstep=1e-9)
    gpu, ref = _pair(base)
    assert np.array_equal(sorted_records(gpu["detp"]), sorted_records(ref["detp"]))
    return base, ref


@pytest.mark.parametrize("outputtype", ["rf", "rfmus", "wltof", "wptof"])
@pytest.mark.parametrize("replaydet", [0, -1])
def test_rf_and_time_weighted_replay(forward, outputtype, replaydet):
    base, fwd = forward
    rep = dict(base, seed=fwd["seeds"], detphotons=fwd["detp"], outputtype=outputtype, replaydet=replaydet,
               omega=2 * math.pi * 150e6)
    gpu, ref = _pair(rep)
    assert np.abs(ref["flux64"]).sum() > 0
    np.testing.assert_allclose(gpu["flux64"], ref["flux64"], rtol=1e-9, atol=1e-12 * np.abs(ref["flux64"]).max())
    np.testing.assert_allclose(gpu["flux"], ref["flux"], rtol=1e-6, atol=1e-9 * np.abs(ref["flux"]).max())
    assert gpu["stat"]["normalizer"] == ref["stat"]["normalizer"]


def test_full_path_fast_mode_statistics():
    """the full path in fast arithmetic (768-lane workgroups, run-time photon
    hand-out): two shared patterns, absorbed fraction and each pattern slot's
    summed fluence against the bit-reproducible run within Monte Carlo error"""
    cfg = workloads.homogeneous_cube_workload(nphoton=400000, srctype="pattern", srcpos=[10, 15, 0],
                                              srcparam1=[40, 0, 0, 4], srcparam2=[0, 30, 0, 3], srcnum=2,
                                              srcpattern=PATS)
    fast = mcxcl_amd.run(dict(cfg, mathmode=0))
    slow = mcxcl_amd.run(dict(cfg, mathmode=1))
    n = cfg["nphoton"]
    assert fast["stat"]["energytot"] == n and slow["stat"]["energytot"] == n
    af, ap = (o["stat"]["energyabs"] / n for o in (fast, slow))
    assert abs(af - ap) <= 6 * np.sqrt(2 * ap * (1 - ap) / n), (af, ap)
    for k in range(2):
        sf = fast["flux"].reshape(2, 60, 60, 60, order="F")[k].sum()
        sp = slow["flux"].reshape(2, 60, 60, 60, order="F")[k].sum()
        assert abs(sf / sp - 1) < 0.01, (k, sf, sp)

"""Device parity of the exit tally per voxel (issaveref > 1, src/mcx_core.cl:852-865):
the sphere-in-cube volume under a background layer, with reflection, against the
oracle with the same seeds and thread map (bit-reproducible mode).  The device
adds with float atomics in arrival order and the oracle sums in double, so the
records agree to float rounding; the fluence and energies are bit-exact as in
any run."""

import numpy as np
import pytest

import mcxcl_amd
import oracle
from mcxcl_amd import workloads

pytestmark = pytest.mark.gpu

NT = 2048


def test_exit_tally_parity():
    vol = workloads.sphere_volume().copy()
    vol[:, :, 0] = 0                                       # background over the tissue
    cfg = workloads.sphere_cube_workload(nphoton=50000, tstep=1e-9, vol=vol, issaveref=2)
    cfg = dict(cfg, autopilot=0, nthread=NT, nblocksize=64, mathmode=1, rawfield=1)
    gpu, ref = mcxcl_amd.run(cfg), oracle.run(cfg, nthread=NT)
    assert ref["detp"].shape == gpu["detp"].shape == (3, 4000000)
    assert ref["stat"]["detected"] == gpu["stat"]["detected"] == 4000000
    assert gpu["stat"]["energyabs"] == ref["stat"]["energyabs"]
    np.testing.assert_allclose(gpu["flux64"], ref["flux64"], rtol=1e-9, atol=1e-12)
    r, d = ref["detp"], gpu["detp"]
    assert (r[0] > 0).sum() > 500                          # voxels of the background layer
    np.testing.assert_allclose(d, r, rtol=2e-5, atol=1e-6 * np.abs(r).max())
    assert not r[2].any() and not d[2].any()               # the last column is never written


def test_exit_tally_fast_build():
    """fast arithmetic: the per-voxel exit weights sum to the bit-reproducible
    build's within Monte Carlo error"""
    vol = workloads.sphere_volume().copy()
    vol[:, :, 0] = 0
    cfg = workloads.sphere_cube_workload(nphoton=400000, tstep=1e-9, vol=vol, issaveref=2)
    fast = mcxcl_amd.run(dict(cfg, mathmode=0))
    slow = mcxcl_amd.run(dict(cfg, mathmode=1))
    sf, sp = (float(o["detp"][0].astype(np.float64).sum()) for o in (fast, slow))
    n = cfg["nphoton"]
    # column 0 holds w (1 + L1): its total is a sum over escaping photons, spread about its mean
    assert abs(sf / sp - 1) < 0.02, (sf, sp)
    assert fast["stat"]["energytot"] == slow["stat"]["energytot"] == n

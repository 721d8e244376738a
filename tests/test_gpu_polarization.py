"""Device parity of polarised transport (src/mcx_core.cl:781-836, :2454-2468,
:2550-2565, :1633-1638, :915-920, :929-959): a 30x30x20 suspension of 0.3 um
spheres (n 1.59 in water, 632.8 nm) under a linearly polarised pencil beam, a
surface detector, reflection on, against the oracle with the same seeds and
thread map (bit-reproducible mode): energies, detected records with their
Stokes columns and Stokes trajectory records bit-exact, the float64 grid within
1e-9 relative.  The fast build agrees within Monte Carlo error."""

import numpy as np
import pytest

import mcxcl_amd
import oracle
from helpers import sorted_records

pytestmark = pytest.mark.gpu

NT = 2048


def _cfg(nphoton, **kw):
    cfg = {"nphoton": nphoton, "vol": np.ones((30, 30, 20), np.uint8), "prop": [[0, 0, 1, 1], [0.005, 0, 0, 1.33]],
           "polprop": [[0.005, 0.3, 0.002, 1.59, 1.33]], "lambda": 632.8, "srcpos": [15, 15, 0],
           "srcdir": [0, 0, 1], "srciquv": [1, 1, 0, 0], "tstart": 0, "tend": 5e-9, "tstep": 1e-9, "seed": 11,
           "detpos": [[15, 15, 0, 6]], "issavedet": 1, "isreflect": 1, "issaveseed": 1}
    cfg.update(kw)
    return cfg


def test_polarised_parity():
    cfg = dict(_cfg(20000), autopilot=0, nthread=NT, nblocksize=64, mathmode=1, rawfield=1)
    gpu, ref = mcxcl_amd.run(cfg), oracle.run(cfg, nthread=NT)
    assert ref["detp"].shape[0] == 10 and ref["detp"].shape[1] > 500
    assert gpu["stat"]["energyabs"] == ref["stat"]["energyabs"]
    np.testing.assert_allclose(gpu["flux64"], ref["flux64"], rtol=1e-9, atol=1e-12)
    assert np.array_equal(sorted_records(gpu["detp"]), sorted_records(ref["detp"]))
    np.testing.assert_array_equal(gpu["prop"], ref["prop"])


def test_polarised_trajectory_records():
    cfg = dict(_cfg(2000, debuglevel="M", istrajstokes=1), autopilot=0, nthread=NT, nblocksize=64, mathmode=1)
    gpu, ref = mcxcl_amd.run(cfg), oracle.run(cfg, nthread=NT)
    assert ref["traj"].shape[0] == 10 and gpu["traj"].shape == ref["traj"].shape
    assert np.array_equal(sorted_records(gpu["traj"]), sorted_records(ref["traj"]))


def test_polarised_fast_build():
    cfg = _cfg(300000)
    n = cfg["nphoton"]
    fast = mcxcl_amd.run(dict(cfg, mathmode=0))
    slow = mcxcl_amd.run(dict(cfg, mathmode=1))
    af, ap = (o["stat"]["energyabs"] / n for o in (fast, slow))
    assert abs(af - ap) <= 6 * np.sqrt(2 * ap * (1 - ap) / n), (af, ap)
    # the fast build samples theta from its marginal and phi given theta (the same
    # distribution as the joint rejection): detected Stokes columns agree in mean
    for row in (-3, -2, -1):
        qf, qp = fast["detp"][row], slow["detp"][row]
        assert abs(qf.mean() - qp.mean()) < 6 * np.sqrt(qf.var() / qf.size + qp.var() / qp.size), row
    nf, ns = fast["detp"].shape[1], slow["detp"].shape[1]
    assert abs(nf - ns) < 6 * np.sqrt(nf + ns)


def test_polarised_2d_parity():
    """a 2D domain: HG with the Mie g, the Stokes vector updated with the sampled
    cosine in the azimuth's place, as the reference's 2D branch does"""
    cfg = dict(_cfg(20000, vol=np.ones((1, 40, 40), np.uint8), srcpos=[0.5, 20, 0], detpos=[[0.5, 20, 0, 8]],
                    issrcfrom0=1),
               autopilot=0, nthread=NT, nblocksize=64, mathmode=1, rawfield=1)
    gpu, ref = mcxcl_amd.run(cfg), oracle.run(cfg, nthread=NT)
    assert ref["detp"].shape[1] > 100
    assert gpu["stat"]["energyabs"] == ref["stat"]["energyabs"]
    np.testing.assert_allclose(gpu["flux64"], ref["flux64"], rtol=1e-9, atol=1e-12)
    assert np.array_equal(sorted_records(gpu["detp"]), sorted_records(ref["detp"]))

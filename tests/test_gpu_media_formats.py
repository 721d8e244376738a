"""Device parity of the packed media formats (src/mcx_core.cl:1079-1193): the
full general path decodes the volume word of every voxel.  Random per-voxel
properties on the 60^3 sphere-in-cube volume with four detectors, against the
oracle with the same seeds and thread map (bit-reproducible mode): energies and
detected-photon records bit-exact, the float64 grid within 1e-9 relative
(atomic summation order only), the float output within 1e-6.  One fast-mode run
per format agrees with the bit-reproducible run within Monte Carlo error.

muamus_half words decode through the same branch as muamus_float words; the
binding has no array form for them (the reference's neither), so they are
covered by the file-reader test and this branch."""

import numpy as np
import pytest

import mcxcl_amd
import oracle
from helpers import sorted_records
from mcxcl_amd import workloads

pytestmark = pytest.mark.gpu

NT = 2048


def _volume(fmt, vol, prop, rng):
    """a packed volume of the sphere workload: per-voxel values around each voxel's
    label properties (void voxels stay void)"""
    lab = vol.astype(np.int64)
    shape = vol.shape
    pert = rng.uniform(0.7, 1.3, shape).astype(np.float32)
    if fmt == "mua_float":
        v = (prop[lab, 0] * pert)[None].astype(np.float32)
        v[0][vol == 0] = np.nan
        return v
    if fmt == "muamus_float":
        v = np.stack([prop[lab, 0] * pert, prop[lab, 1] * (2 - pert)]).astype(np.float32)
        v[:, vol == 0] = np.nan
        return v
    if fmt == "labelplus":
        comp = rng.integers(0, 2, shape).astype(np.float32)                # replace mua or mus
        val = np.where(comp == 0, prop[lab, 0], prop[lab, 1]) * pert
        return np.stack([val, comp, vol.astype(np.float32)]).astype(np.float32)
    if fmt == "asgn_byte":
        b = rng.integers(0, 127, (4,) + shape).astype(np.int8)
        b[:, vol == 0] = 0
        b[0][vol != 0] |= 1                                                 # non-void words
        return b
    if fmt == "muamus_short":
        s = rng.integers(0, 32767, (2,) + shape).astype(np.int16)
        s[:, vol == 0] = 0
        s[0][vol != 0] |= 1
        return s
    raise ValueError(fmt)


def _cfg(fmt, nphoton, seed=0):
    cfg = workloads.sphere_cube_workload(nphoton=nphoton, tstep=1e-9, savedetflag="DXV", issaveseed=1)
    # three rows: the mixing formats interpolate between media 1 and 2
    cfg["prop"] = [[0, 0, 1, 1], [0.005, 1.0, 0.01, 1.37], [0.02, 4.0, 0.5, 1.37]]
    if fmt == "asgn_byte":
        cfg["isreflect"] = 0       # background words decode to medium 1's n: no index step at the faces
    prop = np.asarray(cfg["prop"], np.float32)
    cfg["vol"] = _volume(fmt, cfg["vol"], prop, np.random.default_rng(seed))
    return cfg


FORMATS = ["mua_float", "muamus_float", "labelplus", "asgn_byte", "muamus_short"]


@pytest.mark.parametrize("fmt", FORMATS)
def test_packed_media_parity(fmt):
    cfg = dict(_cfg(fmt, 30000), autopilot=0, nthread=NT, nblocksize=64, mathmode=1, rawfield=1)
    gpu, ref = mcxcl_amd.run(cfg), oracle.run(cfg, nthread=NT)
    assert ref["stat"]["energyabs"] > 0 and ref["detp"].shape[1] > 0
    assert gpu["stat"]["energytot"] == ref["stat"]["energytot"]
    assert gpu["stat"]["energyabs"] == ref["stat"]["energyabs"]
    np.testing.assert_allclose(gpu["flux64"], ref["flux64"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(gpu["flux"], ref["flux"], rtol=1e-6, atol=0)
    assert np.array_equal(sorted_records(gpu["detp"]), sorted_records(ref["detp"]))


@pytest.mark.parametrize("fmt", FORMATS)
def test_packed_media_fast_mode(fmt):
    cfg = _cfg(fmt, 400000, seed=1)
    n = cfg["nphoton"]
    fast = mcxcl_amd.run(dict(cfg, mathmode=0))
    slow = mcxcl_amd.run(dict(cfg, mathmode=1))
    assert fast["stat"]["energytot"] == n and slow["stat"]["energytot"] == n
    af, ap = (o["stat"]["energyabs"] / n for o in (fast, slow))
    assert abs(af - ap) <= 6 * np.sqrt(2 * ap * (1 - ap) / n), (af, ap)
    sf, sp = float(fast["flux"].sum()), float(slow["flux"].sum())
    assert abs(sf / sp - 1) < 0.01, (sf, sp)

"""Device parity of split-voxel media (svmc, src/mcx_core.cl:1229-1348 and the
main-loop hooks at :2638-2701, :2931-2966, :3069-3243): a ball whose surface
voxels are split by tangent planes, against the oracle with the same seeds and
thread map (bit-reproducible mode).  Energies and detected-photon records
bit-exact, the float64 grid within 1e-9 relative (atomic summation order only),
the float output within 1e-6.

Cases: a tissue ball in background with reflection and detectors on its
surface (plane refraction into label 0, its detector capture, the voxel-face
reflection into a split voxel's background side); the same without
reflection (exits on entering label 0); a ball of a second tissue with another
refractive index inside a tissue cube (Fresnel at the plane between two
labels).  The fast build agrees with the bit-reproducible build within Monte
Carlo error, and transmits 1 - R through a tilted plane (Fresnel closed form)."""

import numpy as np
import pytest

import mcxcl_amd
import oracle
from helpers import sorted_records, split_sphere_volume, svmc_bytes

pytestmark = pytest.mark.gpu

NT = 2048
DETS = [[26.5, 20.5, 7.35, 3.0], [20.5, 26.5, 7.35, 3.0], [14.5, 20.5, 7.35, 3.0]]


def _ball_in_air(nphoton, isreflect=1, **over):
    cfg = dict(nphoton=nphoton, vol=split_sphere_volume(40, (20, 20, 20), 14, 1, 0),
               prop=[[0, 0, 1, 1], [0.01, 1.0, 0.5, 1.37]], srcpos=[20.5, 20.5, 0], srcdir=[0, 0, 1],
               tstart=0, tend=5e-9, tstep=5e-10, isreflect=isreflect, issrcfrom0=1, seed=99,
               detpos=DETS, issavedet=1, savedetflag="DSPMXV", issaveseed=1)
    cfg.update(over)
    return cfg


def _ball_in_tissue(nphoton, **over):
    cfg = dict(nphoton=nphoton, vol=split_sphere_volume(40, (20, 20, 20), 12, 2, 1),
               prop=[[0, 0, 1, 1], [0.01, 1.0, 0.5, 1.37], [0.03, 3.0, 0.8, 1.6]],
               srcpos=[20.5, 20.5, 0], srcdir=[0, 0, 1], tstart=0, tend=5e-9, tstep=1e-9, isreflect=1,
               issrcfrom0=1, seed=7)
    cfg.update(over)
    return cfg


def _parity(cfg):
    cfg = dict(cfg, autopilot=0, nthread=NT, nblocksize=64, mathmode=1, rawfield=1)
    gpu, ref = mcxcl_amd.run(cfg), oracle.run(cfg, nthread=NT)
    assert ref["stat"]["energyabs"] > 0
    assert gpu["stat"]["energytot"] == ref["stat"]["energytot"]
    assert gpu["stat"]["energyabs"] == ref["stat"]["energyabs"]
    np.testing.assert_allclose(gpu["flux64"], ref["flux64"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(gpu["flux"], ref["flux"], rtol=1e-6, atol=0)
    return gpu, ref


def test_ball_in_background_with_reflection_and_detectors():
    gpu, ref = _parity(_ball_in_air(20000))
    assert ref["detp"].shape[1] > 100
    assert np.array_equal(sorted_records(gpu["detp"]), sorted_records(ref["detp"]))


def test_ball_in_background_without_reflection():
    gpu, ref = _parity(_ball_in_air(20000, isreflect=0))
    assert ref["detp"].shape[1] > 100
    assert np.array_equal(sorted_records(gpu["detp"]), sorted_records(ref["detp"]))


def test_ball_of_another_tissue():
    _parity(_ball_in_tissue(20000))


@pytest.mark.parametrize("make", [_ball_in_air, _ball_in_tissue])
def test_fast_build_agrees(make):
    cfg = make(300000)
    n = cfg["nphoton"]
    fast = mcxcl_amd.run(dict(cfg, mathmode=0))
    slow = mcxcl_amd.run(dict(cfg, mathmode=1))
    assert fast["stat"]["energytot"] == n and slow["stat"]["energytot"] == n
    af, ap = (o["stat"]["energyabs"] / n for o in (fast, slow))
    assert abs(af - ap) <= 6 * np.sqrt(2 * ap * (1 - ap) / n), (af, ap)
    sf, sp = float(fast["flux"].sum()), float(slow["flux"].sum())
    assert abs(sf / sp - 1) < 0.01, (sf, sp)


def test_fast_build_fresnel_transmission_at_a_tilted_plane():
    """non-scattering pencil beam through a plane tilted 26.5 degrees between
    n = 1 and n = 1.5 in the top layer (absorbing faces): trajectory
    terminations at the far face are the 1 - R of the Fresnel formulas"""
    vol = np.zeros((8, 20, 20, 11), np.uint8)
    vol[0] = 1
    rec = svmc_bytes(1, 2, (51, 128, 255), (191, 128, 255))
    for k in range(8):
        vol[k, :, :, 10] = rec[k]
    N = 200000
    cfg = dict(nphoton=N, vol=vol, prop=[[0, 0, 1, 1], [0, 0, 1, 1.0], [0, 0, 1, 1.5]], srcpos=[10.5, 10.5, 0],
               srcdir=[0, 0, 1], tstart=0, tend=5e-9, tstep=5e-9, isreflect=1, bc="aaaaaa", issrcfrom0=1,
               seed=1234, debuglevel="M", maxjumpdebug=2 * N, mathmode=0)
    tr = mcxcl_amd.run(cfg)["traj"]
    pos, w = tr[1:4], tr[4]
    ends = pos[:, ~((w == 1.0) & (pos[2] == 0.0))]
    assert ends.shape[1] == N
    n = np.array([2 * 191 / 255 - 1, 2 * 128 / 255 - 1, 1.0])
    n /= np.linalg.norm(n)
    ci = n[2]
    ct = np.sqrt(1 - (1 / 1.5) ** 2 * (1 - ci * ci))
    R = 0.5 * (((ci - 1.5 * ct) / (ci + 1.5 * ct)) ** 2 + ((ct - 1.5 * ci) / (ct + 1.5 * ci)) ** 2)
    top = (ends[2] > 10.99).mean()
    assert abs(top - (1 - R)) < 5 * np.sqrt(R * (1 - R) / N), (top, 1 - R)

"""Replay (src/mcx_core.cl:1590-1616 launch, :2845-2858 Jacobian deposit;
src/mcx_utils.c:1355-1466 replay weights) on the CPU oracle.

A replayed photon starts from the RNG state saved when it was detected, so it
retraces its path exactly.  Its Jacobian deposit per voxel is w_i * L_i(v), with
w_i = exp(-sum_j mua_j L_ij) from the recorded partial path lengths.  Summed over
the voxels of one medium this must equal sum_i w_i L_ij, and over all voxels
sum_i w_i L_i: closed forms of the recorded detector data."""

import numpy as np
import pytest

import oracle
from helpers import sorted_records
from mcxcl_amd import workloads

MUA = np.array([0.005, 0.002])   # labels 1 and 2 of the sphere workload


@pytest.fixture(scope="module")
def forward():
    base = workloads.sphere_cube_workload(nphoton=200000, issaveseed=1, savedetflag="DP", tstep=5e-9)
    out = oracle.run(dict(base, autopilot=0, nthread=4096, nblocksize=64), nthread=4096)
    return base, out


def _replay(base, fwd, **extra):
    cfg = dict(base, seed=fwd["seeds"], detphotons=fwd["detp"], outputtype="jacobian", autopilot=0, nthread=4096,
               nblocksize=64, rawfield=1, **extra)
    return oracle.run(cfg, nthread=4096)


def test_replay_jacobian_matches_recorded_paths(forward):
    base, fwd = forward
    detp = fwd["detp"]
    assert detp.shape[1] > 500
    L = detp[1:3].astype(np.float64)
    w = np.exp(-(MUA[:, None] * L).sum(0))
    r = _replay(base, fwd, replaydet=0)
    J = r["flux64"][..., 0]
    assert r["stat"]["nphoton"] == detp.shape[1]
    np.testing.assert_allclose(J.sum(), (w * L.sum(0)).sum(), rtol=1e-6)
    # per medium: the sphere (label 2) and the rest (label 1)
    ii, jj, kk = np.meshgrid(np.arange(60), np.arange(60), np.arange(60), indexing="ij")
    sphere = (ii + 0.5 - 30) ** 2 + (jj + 0.5 - 30) ** 2 + (kk + 0.5 - 30) ** 2 <= 15 ** 2
    np.testing.assert_allclose(J[sphere].sum(), (w * L[1]).sum(), rtol=1e-5)
    np.testing.assert_allclose(J[~sphere].sum(), (w * L[0]).sum(), rtol=1e-6)
    # normalisation: unitinmm / sum of replay weights (src/mcx_host.cpp:1397-1433)
    assert r["stat"]["normalizer"] == pytest.approx(1.0 / w.sum(), rel=1e-6)
    np.testing.assert_allclose(r["flux"][..., 0], (J / w.sum()).astype(np.float32), rtol=2e-6, atol=1e-12)


def test_replay_per_detector_slices(forward):
    """replaydet = -1: one Jacobian per detector, stacked after the gates"""
    base, fwd = forward
    detp = fwd["detp"]
    det = detp[0].astype(np.int64) & 0xFFFF
    L = detp[1:3].astype(np.float64)
    w = np.exp(-(MUA[:, None] * L).sum(0))
    r = _replay(base, fwd, replaydet=-1)
    J = r["flux64"]
    assert J.shape[-1] == 4
    for d in range(1, 5):
        sel = det == d
        np.testing.assert_allclose(J[..., d - 1].sum(), (w[sel] * L[:, sel].sum(0)).sum(), rtol=1e-6)


def test_replay_single_detector(forward):
    """replaydet = d: only photons of detector d are replayed"""
    base, fwd = forward
    detp = fwd["detp"]
    det = detp[0].astype(np.int64) & 0xFFFF
    L = detp[1:3].astype(np.float64)
    w = np.exp(-(MUA[:, None] * L).sum(0))
    r = _replay(base, fwd, replaydet=2)
    sel = det == 2
    np.testing.assert_allclose(r["flux64"].sum(), (w[sel] * L[:, sel].sum(0)).sum(), rtol=1e-6)


def test_label_jacobian_matches_finite_difference(forward):
    """d S / d mua_1 by central differences over forward runs with the same seed
    (absorption does not change the paths) equals -sum of the Jacobian over the
    voxels of label 1"""
    base, fwd = forward
    r = _replay(base, fwd, replaydet=0)
    ii, jj, kk = np.meshgrid(np.arange(60), np.arange(60), np.arange(60), indexing="ij")
    sphere = (ii + 0.5 - 30) ** 2 + (jj + 0.5 - 30) ** 2 + (kk + 0.5 - 30) ** 2 <= 15 ** 2
    grad = -r["flux64"][..., 0][~sphere].sum()
    h = 2e-4
    sig = []
    for mua1 in (0.005 + h, 0.005 - h):
        prop = [[0, 0, 1, 1], [mua1, 1.0, 0.01, 1.37], [0.002, 5.0, 0.9, 1.0]]
        out = oracle.run(dict(base, prop=prop, autopilot=0, nthread=4096, nblocksize=64), nthread=4096)
        # same photons, same paths (records come back in detection order, so sort)
        assert np.array_equal(sorted_records(out["detp"]), sorted_records(fwd["detp"]))
        Lp = out["detp"][1:3].astype(np.float64)
        sig.append(np.exp(-(np.array([mua1, 0.002])[:, None] * Lp).sum(0)).sum())
    fd = (sig[0] - sig[1]) / (2 * h)
    assert abs(fd - grad) / abs(grad) < 1e-3

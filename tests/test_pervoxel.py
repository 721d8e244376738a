"""Per-voxel absorption/scattering media and the replay adjoint on the CPU oracle.

With per-voxel media a voxel of non-zero label takes mua and mus from the
per-voxel array and g and n from its label.  Detected photons report their exit
weight w_i = exp(-sum_v mua_v L_iv); replaying them with those weights gives the
absorption Jacobian J(v) = sum_i w_i L_iv = -dS/dmua_v, S = sum_i w_i.  Absorption
does not move photons, so forward runs with the same seed and a perturbed mua
detect the same photons along the same paths: central differences of S are exact
up to O(h^2)."""

import numpy as np
import pytest

import oracle
from helpers import sorted_records
from mcxcl_amd import workloads

NT = 4096


def _run(cfg, **kw):
    return oracle.run(dict(cfg, autopilot=0, nthread=NT, nblocksize=64, **kw), nthread=NT)


def test_pervoxel_media_equal_to_labels_reproduce_the_label_run():
    base = workloads.sphere_cube_workload(nphoton=50000, issaveseed=1, savedetflag="DPX", tstep=5e-9)
    lab = _run(base, rawfield=1)
    vol = base["vol"]
    prop = np.asarray(base["prop"], np.float32)
    mm = np.stack([prop[vol, 0], prop[vol, 1]]).astype(np.float32)
    per = _run(dict(base, muamus=mm), rawfield=1)
    assert per["stat"]["energyabs"] == lab["stat"]["energyabs"]
    np.testing.assert_allclose(per["flux64"], lab["flux64"], rtol=1e-12, atol=1e-15)
    assert np.array_equal(sorted_records(per["detp"]), sorted_records(lab["detp"]))


def _field(seed=5):
    """a smooth random mua/mus field on the 60^3 sphere workload"""
    rng = np.random.default_rng(seed)
    base = workloads.sphere_cube_workload(nphoton=100000, issaveseed=1, savedetflag="DP", tstep=5e-9)
    vol = base["vol"]
    prop = np.asarray(base["prop"], np.float32)
    ii, jj, kk = np.meshgrid(np.arange(60), np.arange(60), np.arange(60), indexing="ij")
    pert = 1 + 0.3 * np.sin(ii / 7.0 + rng.uniform(0, 6)) * np.cos(jj / 9.0 + rng.uniform(0, 6)) * \
        np.sin(kk / 5.0 + rng.uniform(0, 6))
    mua = (prop[vol, 0] * pert).astype(np.float32)
    mus = (prop[vol, 1] * (2 - pert)).astype(np.float32)
    return base, np.stack([mua, mus])


def test_exit_weights_match_pathlength_weights():
    """label media: the recorded exit weight equals exp(-sum_j mua_j L_j) from the
    partial path lengths (the reference's host-side detector weight)"""
    base = workloads.sphere_cube_workload(nphoton=100000, issaveseed=1, savedetflag="DP", tstep=5e-9)
    out = _run(base, savedetweight=1)
    L = out["detp"][1:3].astype(np.float64)
    w = np.exp(-(np.array([0.005, 0.002])[:, None] * L).sum(0))
    np.testing.assert_allclose(out["detw"], w, rtol=2e-5)


def test_pervoxel_replay_jacobian_and_finite_differences():
    base, mm = _field()
    fwd = _run(dict(base, muamus=mm), savedetweight=1)
    detp, w = fwd["detp"], fwd["detw"].astype(np.float64)
    assert detp.shape[1] > 300
    rep = _run(dict(base, muamus=mm, seed=fwd["seeds"], detphotons=detp, replayweight=fwd["detw"],
                    outputtype="jacobian", replaydet=0), rawfield=1)
    J = rep["flux64"][..., 0]
    # total: sum_i w_i L_i from the recorded path lengths
    np.testing.assert_allclose(J.sum(), (w * detp[1:3].astype(np.float64).sum(0)).sum(), rtol=1e-5)
    # finite differences at the 8 voxels of largest sensitivity and 4 random ones
    order = np.argsort(J.ravel())[::-1]
    rng = np.random.default_rng(1)
    picks = list(order[:8]) + list(rng.choice(order[:2000], 4, replace=False))
    S0 = w.sum()
    for flat in picks:
        v = np.unravel_index(flat, J.shape)
        h = 0.05 * mm[0][v]
        sig = []
        for s in (1, -1):
            m2 = mm.copy()
            m2[0][v] += s * h
            o = _run(dict(base, muamus=m2), savedetweight=1)
            sig.append(o["detw"].astype(np.float64).sum())
        fd = (sig[0] - sig[1]) / (2 * h)
        assert abs(fd + J[v]) <= 2e-3 * J[v] + 1e-6 * S0, (v, fd, -J[v])

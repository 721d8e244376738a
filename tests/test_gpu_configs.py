"""The layered, batched and inverse-step workloads at their full sizes on the
device (fast arithmetic, as the benchmark runs them).

Layered forward + adjoint: the 200^3 four-layer volume, 1e8 photons, 8
detectors.  The replay Jacobian summed over each layer equals the recorded
sum_i w_i L_i(layer) of the detected photons (src/mcx_utils.c:1355-1421
weights, src/mcx_core.cl:2845-2858 deposits), and the gradient of the summed
signals with respect to per-voxel mua matches central differences with
common random numbers at voxels drawn uniformly from the sensitive region.

Batched forward: all 256 instances of 64^3 in one persistent launch.  Every
instance launches exactly its photons, and instances compared with oracle runs
of a sample of their photons agree within Monte Carlo error (absorbed fraction
and normalised fluence sum, the oracle's float64 tally).

Inverse step: 128^3 per-voxel media, 16 sources x 16 detectors, 1e7 photons
per source in one launch.  Every pair within 50 mm sees light, the all-pairs gradients
of one signal match central differences, and the folded backward of a
weighted loss equals the weighted sum of the all-pairs gradients."""

import numpy as np
import pytest
import torch

import mcxcl_amd
import oracle
from mcxcl_amd import autograd, pmcxcl, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture
def fixed_map(monkeypatch):
    # every forward run launches the same photons on the same lanes
    monkeypatch.setenv("MCXHIP_DYNAMIC", "0")


def test_layered_200_forward_and_adjoint(fixed_map):
    cfg = workloads.layered_workload(nphoton=100000000, mathmode=0)
    vol = cfg["vol"]
    prop = np.asarray(cfg["prop"], np.float32)
    mua0, mus0 = prop[vol, 0].astype(np.float32), prop[vol, 1].astype(np.float32)

    # replay Jacobian per layer against the recorded path lengths
    fwd = mcxcl_amd.run(dict(cfg, issaveseed=1, issave2pt=0))
    detp = fwd["detp"]
    assert detp.shape[1] > 100000
    L = detp[1:5].astype(np.float64)
    w = np.exp(-(prop[1:5, 0].astype(np.float64)[:, None] * L).sum(0))
    rep = mcxcl_amd.run(dict(cfg, seed=fwd["seeds"], detphotons=detp, outputtype="jacobian", replaydet=0,
                             rawfield=1, rawonly=1))
    J = rep["flux64"][..., 0]
    for lab in range(1, 5):
        np.testing.assert_allclose(J[vol == lab].sum(), (w * L[lab - 1]).sum(), rtol=2e-5)

    # autograd gradient against central differences
    mua = torch.tensor(mua0, requires_grad=True)
    autograd.detector_signals(mua, torch.tensor(mus0), cfg).sum().backward()
    g = mua.grad.numpy()
    sensitive = np.flatnonzero(np.abs(g.ravel()) >= 1e-4 * np.abs(g).max())
    assert sensitive.size > 1000
    rng = np.random.default_rng(7)
    for flat in rng.choice(sensitive, 4, replace=False):
        v = np.unravel_index(flat, g.shape)
        h = mua0[v]
        sig = []
        for s in (1, -1):
            m = mua0.copy()
            m[v] += s * h
            sig.append(float(autograd.detector_signals(torch.tensor(m), torch.tensor(mus0), cfg).sum()))
        fd = (sig[0] - sig[1]) / (2 * h)
        assert abs(fd - g[v]) <= 0.02 * abs(g[v]), (v, fd, g[v])


def test_batched_256x64_every_instance_and_oracle_samples():
    cfgs = workloads.batched_workload(ninst=256)
    outs, info = pmcxcl.run_batch(cfgs, usegraph=1)
    assert len(outs) == 256
    for out, c in zip(outs, cfgs):
        assert out["stat"]["energytot"] == c["nphoton"], c["session"]
    # four instances, one of each source type plus the one with the most photons,
    # against the oracle on 20000 of their photons
    kinds = {}
    for i, c in enumerate(cfgs):
        kinds.setdefault(c["srctype"], i)
    picks = sorted(set(kinds.values()) | {int(np.argmax([c["nphoton"] for c in cfgs]))})
    assert len(picks) >= 4
    for i in picks:
        c, out = cfgs[i], outs[i]
        n_gpu, n_ref = c["nphoton"], 20000
        ref = oracle.run(dict(c, nphoton=n_ref, autopilot=0, nthread=2048, nblocksize=64), nthread=2048)
        a_gpu = out["stat"]["energyabs"] / n_gpu
        a_ref = ref["stat"]["energyabs"] / ref["stat"]["energytot"]
        se = np.sqrt(a_ref * (1 - a_ref) * (1.0 / n_ref + 1.0 / n_gpu))
        assert abs(a_gpu - a_ref) < 5 * se + 1e-4, (c["session"], a_gpu, a_ref, se)
        # normalised fluence sums: both are fluence rate per launched photon
        f_gpu, f_ref = float(out["flux"].sum()), float(ref["flux"].sum())
        assert abs(f_gpu / f_ref - 1) < 0.05, (c["session"], f_gpu, f_ref)


def test_inverse_step_128_16x16(fixed_map):
    cfg, mua0, mus0 = workloads.inverse_workload()
    S, dmua, dmus = autograd.signal_jacobians(mua0, mus0, cfg)
    assert S.shape == (16, 16) and (S >= 0).all() and np.isfinite(S).all()
    # every pair of the two nearest spacings (20.1 and 45.0 mm) sees light (pairs
    # across the 128 mm face may see none at 1e7 photons per source); the signal
    # falls with the source-detector distance
    src = np.asarray(cfg["srcpos"], np.float64)[:, :2]
    det = np.asarray(cfg["detpos"], np.float64)[:, :2]
    dist = np.linalg.norm(src[:, None, :] - det[None, :, :], axis=-1)
    near, mid = dist < 25, (dist > 40) & (dist < 50)
    assert near.sum() >= 16 and mid.sum() >= 16
    assert (S[dist < 50] > 0).all()
    assert S[near].mean() > 30 * S[mid].mean()
    assert dmua.shape == (16, 16, 128, 128, 128) and np.isfinite(dmus).all()
    # the folded backward of a weighted loss is the weighted sum of the all-pairs gradients
    wts = np.random.default_rng(3).uniform(0.5, 1.5, (16, 16))
    mua = torch.tensor(mua0, requires_grad=True)
    (autograd.detector_signals(mua, torch.tensor(mus0), cfg) * torch.tensor(wts)).sum().backward()
    ga = np.tensordot(wts, dmua.astype(np.float64), axes=([0, 1], [0, 1]))
    np.testing.assert_allclose(mua.grad.numpy(), ga, rtol=1e-3, atol=1e-5 * np.abs(ga).max())
    # one signal against central differences at two sensitive voxels
    s_, d_ = 5, 6
    g = dmua[s_, d_]
    sensitive = np.flatnonzero(np.abs(g.ravel()) >= 1e-4 * np.abs(g).max())
    for flat in np.random.default_rng(11).choice(sensitive, 2, replace=False):
        v = np.unravel_index(flat, g.shape)
        h = mua0[v]
        sig = []
        for sgn in (1, -1):
            m = mua0.copy()
            m[v] += sgn * h
            sig.append(float(autograd.detector_signals(torch.tensor(m), torch.tensor(mus0), cfg)[s_, d_]))
        fd = (sig[0] - sig[1]) / (2 * h)
        assert abs(fd - g[v]) <= 0.02 * abs(g[v]), (v, fd, g[v])

"""mcxcl_amd.autograd: gradients of detector signals by replay.

CPU tests run the wrapper with the oracle standing in for the device, so they
check the wrapper's assembly of J and WP into gradients; the GPU test at the end
runs the same checks on the device.  The absorption gradient is compared with
central finite differences (exact common random numbers: absorption does not
move photons).  The scattering gradient WP/mus - J is the score of the path
density; its two deposits are pinned against the recorded per-photon scattering
counts and path lengths (a finite difference in mus changes every later random
number of a lane, so it is no common-random-number check)."""

import numpy as np
import pytest
import torch

import oracle
from mcxcl_amd import autograd, workloads

NT = 4096


def _oracle_run(cfg):
    return oracle.run(dict(cfg, autopilot=0, nthread=NT, nblocksize=64), nthread=NT)


@pytest.fixture
def oracle_engine(monkeypatch):
    monkeypatch.setattr(autograd, "_run", _oracle_run)


def _fields(base, seed=3):
    vol = base["vol"]
    prop = np.asarray(base["prop"], np.float32)
    rng = np.random.default_rng(seed)
    ii, jj, kk = np.meshgrid(np.arange(60), np.arange(60), np.arange(60), indexing="ij")
    pert = 1 + 0.25 * np.sin(ii / 6.0 + rng.uniform(0, 6)) * np.cos(kk / 8.0 + rng.uniform(0, 6))
    return (prop[vol, 0] * pert).astype(np.float32), (prop[vol, 1] * (2 - pert)).astype(np.float32)


def test_mua_gradient_matches_finite_differences(oracle_engine):
    base = workloads.sphere_cube_workload(nphoton=100000, tstep=5e-9)
    mua0, mus0 = _fields(base)
    mua = torch.tensor(mua0, requires_grad=True)
    mus = torch.tensor(mus0, requires_grad=True)
    wts = torch.tensor([1.0, -0.5, 2.0, 0.25], dtype=torch.float64)
    S = autograd.detector_signals(mua, mus, base)
    (S * wts).sum().backward()
    g = mua.grad.numpy()
    order = np.argsort(np.abs(g).ravel())[::-1]
    for flat in list(order[:6]) + list(order[500:502]):
        v = np.unravel_index(flat, g.shape)
        h = 0.05 * mua0[v]
        sig = []
        for s in (1, -1):
            m = mua0.copy()
            m[v] += s * h
            sig.append(float((autograd.detector_signals(torch.tensor(m), torch.tensor(mus0), base) * wts).detach().sum()))
        fd = (sig[0] - sig[1]) / (2 * h)
        assert abs(fd - g[v]) <= 2e-3 * abs(g[v]) + 1e-4, (v, fd, g[v])


def test_scattering_deposits_match_recorded_counts(oracle_engine):
    """sum over the voxels of each label of WP = sum_i w_i n_i(label) and of J =
    sum_i w_i L_i(label), from the recorded records (savedetflag DSP)"""
    base = workloads.sphere_cube_workload(nphoton=100000, tstep=5e-9)
    mua0, mus0 = _fields(base)
    mm = np.stack([mua0, mus0])
    fwd = _oracle_run(dict(base, muamus=mm, issaveseed=1, savedetweight=1, savedetflag="DSP"))
    detp, w = fwd["detp"], fwd["detw"].astype(np.float64)
    nscat = detp[1:3].view(np.uint32).astype(np.float64)
    L = detp[3:5].astype(np.float64)
    rep = dict(base, muamus=mm, seed=fwd["seeds"], detphotons=detp, replayweight=fwd["detw"], replaydet=0,
               rawfield=1, savedetflag="DSP")
    wp = _oracle_run(dict(rep, outputtype="wp"))["flux64"][..., 0]
    jac = _oracle_run(dict(rep, outputtype="jacobian"))["flux64"][..., 0]
    vol = base["vol"]
    for lab in (1, 2):
        np.testing.assert_allclose(wp[vol == lab].sum(), (w * nscat[lab - 1]).sum(), rtol=1e-6)
        np.testing.assert_allclose(jac[vol == lab].sum(), (w * L[lab - 1]).sum(), rtol=1e-5)


def test_gradient_of_nothing_detected_is_zero(oracle_engine):
    base = workloads.homogeneous_cube_workload(nphoton=2000, detpos=[[5, 5, 59, 1]], issavedet=1,
                                               savedetflag="DP", tend=1e-11, tstep=1e-11)
    vol = base["vol"]
    prop = np.asarray(base["prop"], np.float32)
    mua = torch.tensor(prop[vol, 0], requires_grad=True)
    mus = torch.tensor(prop[vol, 1], requires_grad=True)
    S = autograd.detector_signals(mua, mus, base)
    assert float(S.detach().sum()) == 0.0
    S.sum().backward()
    assert float(mua.grad.abs().sum()) == 0.0 and float(mus.grad.abs().sum()) == 0.0


@pytest.mark.gpu
def test_mua_gradient_on_device_matches_finite_differences():
    """config-2 style check on the device: central differences at 16 voxels,
    at most 2% relative error"""
    # bit-reproducible mode: every forward run launches the same photons on the same lanes
    base = workloads.sphere_cube_workload(nphoton=2000000, tstep=5e-9, mathmode=1)
    mua0, mus0 = _fields(base)
    mua = torch.tensor(mua0, requires_grad=True)
    mus = torch.tensor(mus0, requires_grad=True)
    S = autograd.detector_signals(mua, mus, base)
    S.sum().backward()
    g = mua.grad.numpy()
    order = np.argsort(np.abs(g).ravel())[::-1]
    rng = np.random.default_rng(0)
    picks = list(order[:8]) + list(rng.choice(order[:5000], 8, replace=False))
    for flat in picks:
        v = np.unravel_index(flat, g.shape)
        h = 0.1 * mua0[v]
        sig = []
        for s in (1, -1):
            m = mua0.copy()
            m[v] += s * h
            sig.append(float(autograd.detector_signals(torch.tensor(m), torch.tensor(mus0), base).detach().sum()))
        fd = (sig[0] - sig[1]) / (2 * h)
        assert abs(fd - g[v]) <= 0.02 * abs(g[v]), (v, fd, g[v])

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
    out = oracle.run(dict(cfg, autopilot=0, nthread=NT, nblocksize=64, rawfield=1), nthread=NT)
    if cfg.get("rawfield32"):
        # the device converts its float64 accumulator to float32 the same way
        out["flux32"] = np.asfortranarray(out["flux64"].astype(np.float32))
    return out


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


def _multisource_case(nphoton=30000):
    """3 pencil sources and 2 detectors on a 40^3 two-layer volume with per-voxel media"""
    vol = np.ones((40, 40, 40), np.uint8)
    vol[:, :, 15:] = 2
    cfg = dict(nphoton=nphoton, vol=vol, prop=[[0, 0, 1, 1], [0.01, 3.0, 0.8, 1.37], [0.02, 5.0, 0.8, 1.37]],
               srcpos=[[12, 20, 0], [20, 20, 0], [28, 20, 0]], srcdir=[[0, 0, 1, 0]] * 3, detpos=[[16, 20, 0, 2],
               [24, 20, 0, 2]], tstart=0, tend=5e-9, tstep=5e-9, issrcfrom0=1, isreflect=1, seed=77,
               savedetflag="DP", maxdetphoton=200000)
    prop = np.asarray(cfg["prop"], np.float32)
    rng = np.random.default_rng(1)
    pert = 1 + 0.2 * rng.uniform(-1, 1, vol.shape).astype(np.float32)
    return cfg, (prop[vol, 0] * pert).astype(np.float32), (prop[vol, 1] * (2 - pert)).astype(np.float32)


def test_multisource_signals_are_per_source_and_detector(oracle_engine):
    cfg, mua0, mus0 = _multisource_case()
    S = autograd.detector_signals(torch.tensor(mua0), torch.tensor(mus0), cfg)
    assert tuple(S.shape) == (3, 2)
    # the source beside each detector gives it the most light
    s = S.numpy()
    assert s[0, 0] > s[2, 0] and s[2, 1] > s[0, 1] and (s > 0).all()


def test_multisource_gradient_of_one_pair_is_its_replay_jacobian(oracle_engine):
    """d S_sd / d mua = -J_sd: the replay Jacobian of the photons of source s
    seen by detector d (the reference's replay with srcid=s, replaydet=d,
    src/mcx_utils.c:1355-1421, src/mcx_core.cl:1602-1616)"""
    cfg, mua0, mus0 = _multisource_case()
    for s_, d_ in [(0, 1), (2, 0)]:
        mua = torch.tensor(mua0, requires_grad=True)
        S = autograd.detector_signals(mua, torch.tensor(mus0), cfg)
        S[s_, d_].backward()
        fwd = _oracle_run(dict(cfg, muamus=np.stack([mua0, mus0]), issavedet=1, issaveseed=1, savedetweight=1,
                               issave2pt=0, srcid=0))
        rep = dict(cfg, muamus=np.stack([mua0, mus0]), seed=fwd["seeds"], detphotons=fwd["detp"],
                   replayweight=fwd["detw"], replaydet=d_ + 1, srcid=s_ + 1, rawfield=1, outputtype="jacobian")
        jac = _oracle_run(rep)["flux64"][..., 0]
        np.testing.assert_allclose(mua.grad.numpy(), -jac, rtol=1e-5, atol=1e-6 * np.abs(jac).max())


def test_multisource_mua_gradient_matches_finite_differences(oracle_engine):
    cfg, mua0, mus0 = _multisource_case(nphoton=20000)
    wts = torch.tensor([[1.0, -0.5], [2.0, 0.3], [-1.0, 0.7]], dtype=torch.float64)
    mua = torch.tensor(mua0, requires_grad=True)
    S = autograd.detector_signals(mua, torch.tensor(mus0), cfg)
    (S * wts).sum().backward()
    g = mua.grad.numpy()
    order = np.argsort(np.abs(g).ravel())[::-1]
    for flat in list(order[:4]) + list(order[300:302]):
        v = np.unravel_index(flat, g.shape)
        h = 0.05 * mua0[v]
        sig = []
        for sgn in (1, -1):
            m = mua0.copy()
            m[v] += sgn * h
            sig.append(float((autograd.detector_signals(torch.tensor(m), torch.tensor(mus0), cfg) * wts).sum()))
        fd = (sig[0] - sig[1]) / (2 * h)
        assert abs(fd - g[v]) <= 2e-3 * abs(g[v]) + 1e-4, (v, fd, g[v])


def test_signal_jacobians_match_the_folded_backward(oracle_engine):
    """every (source, detector) slice of the all-pairs replay equals the
    gradient of that one signal through the folded backward, and the slices
    weighted by an upstream gradient sum to the folded result"""
    cfg, mua0, mus0 = _multisource_case(nphoton=20000)
    S, dmua, dmus = autograd.signal_jacobians(mua0, mus0, cfg)
    assert dmua.shape == (3, 2, 40, 40, 40) and dmus.shape == dmua.shape
    S2 = autograd.detector_signals(torch.tensor(mua0), torch.tensor(mus0), cfg).numpy()
    np.testing.assert_allclose(S, S2, rtol=1e-12)
    wts = np.array([[1.0, -0.5], [2.0, 0.3], [-1.0, 0.7]])
    mua = torch.tensor(mua0, requires_grad=True)
    mus = torch.tensor(mus0, requires_grad=True)
    (autograd.detector_signals(mua, mus, cfg) * torch.tensor(wts)).sum().backward()
    ga = np.tensordot(wts, dmua.astype(np.float64), axes=([0, 1], [0, 1]))
    gs = np.tensordot(wts, dmus.astype(np.float64), axes=([0, 1], [0, 1]))
    np.testing.assert_allclose(mua.grad.numpy(), ga, rtol=1e-4, atol=1e-5 * np.abs(ga).max())
    np.testing.assert_allclose(mus.grad.numpy(), gs, rtol=1e-4, atol=1e-5 * np.abs(gs).max())


@pytest.mark.gpu
def test_multisource_jacobians_on_device_match_finite_differences():
    """the inverse-step workload at reduced size on the device: all-pairs
    gradients against central differences of single signals at 6 voxels each
    (bit-reproducible mode: every forward run launches the same photons)"""
    cfg, mua0, mus0 = workloads.inverse_workload(dim=48, photons_per_source=100000, nside=2, top=12, mathmode=1)
    S, dmua, dmus = autograd.signal_jacobians(mua0, mus0, cfg)
    assert S.shape == (4, 4) and (S > 0).all() and dmua.shape == (4, 4, 48, 48, 48)
    rng = np.random.default_rng(3)
    for s_, d_ in [(0, 0), (3, 1)]:
        g = dmua[s_, d_]
        order = np.argsort(np.abs(g).ravel())[::-1]
        for flat in list(order[:3]) + list(rng.choice(order[:2000], 3, replace=False)):
            v = np.unravel_index(flat, g.shape)
            h = 0.1 * mua0[v]
            sig = []
            for sgn in (1, -1):
                m = mua0.copy()
                m[v] += sgn * h
                sig.append(float(autograd.detector_signals(torch.tensor(m), torch.tensor(mus0), cfg)[s_, d_]))
            fd = (sig[0] - sig[1]) / (2 * h)
            assert abs(fd - g[v]) <= 0.02 * abs(g[v]), (s_, d_, v, fd, g[v])


def _mus_scaling_check(base, label, h, seeds, photons):
    """d S / d a for mus -> mus (1 + a) over the voxels of one label: the score
    gradient sum_v mus_v dS/dmus_v against central differences S(+h), S(-h) on
    the same seeds (common random numbers up to the first changed flight).
    Returns per-seed (gradient, difference, sum of the WP term)."""
    vol = base["vol"]
    prop = np.asarray(base["prop"], np.float32)
    mua0 = prop[vol, 0].astype(np.float32)
    mus0 = prop[vol, 1].astype(np.float32)
    sel = vol == label
    rows = []
    for seed in seeds:
        cfg = dict(base, seed=int(seed), nphoton=int(photons))
        mus = torch.tensor(mus0, requires_grad=True)
        mua = torch.tensor(mua0, requires_grad=True)
        autograd.detector_signals(mua, mus, cfg).sum().backward()
        grad = float((mus0[sel].astype(np.float64) * mus.grad.numpy()[sel]).sum())
        wp_term = float((mus0[sel].astype(np.float64) * (mus.grad.numpy()[sel] - mua.grad.numpy()[sel])).sum())
        sig = []
        for s in (1, -1):
            m = mus0.copy()
            m[sel] *= np.float32(1 + s * h)
            sig.append(float(autograd.detector_signals(torch.tensor(mua0), torch.tensor(m), cfg).sum()))
        rows.append((grad, (sig[0] - sig[1]) / (2 * h), wp_term))
    return np.array(rows)


def _assert_mus_scaling(rows):
    g, fd, wp = rows[:, 0], rows[:, 1], rows[:, 2]
    d = fd - g
    se = d.std(ddof=1) / np.sqrt(len(d))
    assert abs(d.mean()) < 4.5 * se, (g.mean(), fd.mean(), se)
    # the check has the power to see either score term dropped
    assert 10 * se < abs(wp.mean()), (se, wp.mean())


def test_mus_gradient_matches_layerwise_finite_differences(oracle_engine):
    """dS/dmus = WP/mus - J: scaling mus of the whole background label by
    (1 +- 10%) on 6 seeds; the mean difference between the score gradient and
    the central difference is zero within 4.5 standard errors, and that error
    is below a tenth of the WP term (dropping WP/mus or -J fails)"""
    base = workloads.sphere_cube_workload(nphoton=100000, tstep=5e-9)
    _assert_mus_scaling(_mus_scaling_check(base, 1, 0.1, 1000 + 17 * np.arange(6), 100000))


@pytest.mark.gpu
def test_mus_gradient_on_device_matches_layerwise_finite_differences(monkeypatch):
    """the same check on the device at 2e6 photons per seed, 8 seeds, for the
    background and the sphere label (fast arithmetic, fixed photon-per-lane map
    so each seed's three runs share their random numbers)"""
    monkeypatch.setenv("MCXHIP_DYNAMIC", "0")
    base = workloads.sphere_cube_workload(nphoton=2000000, tstep=5e-9, mathmode=0)
    for label in (1, 2):
        _assert_mus_scaling(_mus_scaling_check(base, label, 0.1, 500 + 31 * np.arange(8), 2000000))

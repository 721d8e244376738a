"""The benchmarked fast build against the bit-reproducible build on the benchmark
workload (60^3 cube + r=15 sphere, 10 gates over 0-5 ns, 4 r=1 mm detectors,
D/P/X records; the reference's cubesph60b with 10 gates).

The bit-reproducible build equals the CPU oracle bit for bit (test_gpu_parity.py),
so it stands in for the oracle at sizes the oracle cannot run in seconds.  The
fast build uses hardware transcendentals, run-time photon hand-out and a float32
LDS deposit box, so it is compared statistically: each build runs K independent
seeds, and every quantity is compared by Welch's z = (mean_f - mean_p) /
sqrt(s_f^2/K + s_p^2/K), with sigma from the batch means.  With K = 8 per side the
statistic follows a t distribution with about 14 degrees of freedom; |z| < 5.5
has a false-alarm rate near 1e-4 per quantity.

Quantities: absorbed fraction, per-gate totals, the fluence summed over the
region around the source (where the fast build's LDS box collects its deposits)
and over the rest of the grid per gate, hot-voxel fluence (mean z^2 over the
hottest voxels, E = nu/(nu-2) = 1.17 for no difference), detected photons per
launched photon per detector and their mean partial path lengths per medium.

Two more runs put the float32 box under load: one of 1e8 photons (the benchmark
itself, ~390k photons per workgroup) and one of 1e7 photons on 8 workgroups
(1.25M photons per workgroup), each against the parity batch means with the
variance scaled by the photon ratio."""

import os

import numpy as np
import pytest

import mcxcl_amd
from mcxcl_amd import workloads

pytestmark = pytest.mark.gpu

K = 8                 # independent seeds per build
NB = 1250000          # photons per batch (K * NB = 1e7 per build)
ZMAX = 5.5
BOX = (slice(17, 42), slice(17, 42), slice(0, 25))   # 25^3 voxels around the source (29,29,0)


def _summary(out):
    """the compared quantities of one run, as a flat dict of float64 arrays"""
    flux = out["flux"].astype(np.float64)          # [60,60,60,10], normalised fluence rate
    st = out["stat"]
    nphot = st["energytot"]
    det = out["detp"]
    detid = det[0].astype(np.int64) & 0xFFFF
    ppath = det[1:3].astype(np.float64)
    s = {
        "absorbed": np.array([st["energyabs"] / st["energytot"]]),
        "gate_total": flux.sum(axis=(0, 1, 2)),
        "box_total": flux[BOX].sum(axis=(0, 1, 2))[:3],
        "rest_total": (flux.sum(axis=(0, 1, 2)) - flux[BOX].sum(axis=(0, 1, 2)))[:3],
        "detected": np.array([(detid == d).sum() / nphot for d in range(1, 5)]),
        "ppath": np.concatenate([ppath[:, detid == d].mean(axis=1) for d in range(1, 5)]),
    }
    return s, flux[..., :2]


def _batches(mathmode, seed0):
    sums, hot = [], []
    for k in range(K):
        cfg = workloads.sphere_cube_workload(nphoton=NB, seed=seed0 + 7919 * k, mathmode=mathmode)
        s, f = _summary(mcxcl_amd.run(cfg))
        sums.append(s)
        hot.append(f)
    keys = sums[0].keys()
    return {k: np.stack([s[k] for s in sums]) for k in keys}, np.stack(hot)


@pytest.fixture(scope="module")
def batches():
    print("fast-mode statistics: %d parity and %d fast batches of %d photons" % (K, K, NB), flush=True)
    par, par_hot = _batches(1, 1648335518)
    fast, fast_hot = _batches(0, 20251205)
    return par, par_hot, fast, fast_hot


def _welch(a, b):
    ma, mb = a.mean(0), b.mean(0)
    se = np.sqrt(a.var(0, ddof=1) / a.shape[0] + b.var(0, ddof=1) / b.shape[0])
    return (mb - ma) / np.where(se > 0, se, np.inf)


def _hot_mask(par_hot):
    """the 2000 hottest voxels of the first two gates, by the parity mean"""
    m = par_hot.mean(0)
    thr = np.sort(m.ravel())[-2000]
    return m >= thr


def test_fast_matches_parity_batches(batches):
    par, par_hot, fast, fast_hot = batches
    for key in par:
        z = _welch(par[key], fast[key])
        print("%-10s max|z| %.2f  fast/parity %s" % (key, np.abs(z).max(),
                                                     np.array2string(fast[key].mean(0) / par[key].mean(0),
                                                                     precision=5, max_line_width=200)), flush=True)
        assert np.all(np.abs(z) < ZMAX), (key, z)
    mask = _hot_mask(par_hot)
    z = _welch(par_hot[:, mask], fast_hot[:, mask])
    print("hot voxels: mean z %.3f, mean z^2 %.3f" % (z.mean(), (z * z).mean()), flush=True)
    assert (z * z).mean() < 1.6
    # every photon is launched exactly once in both builds
    assert abs(par["absorbed"].mean() - 0.2702) < 0.002


def _single_vs_batches(out, nphoton, par, par_hot):
    s, f = _summary(out)
    scale = NB / float(nphoton)          # variance of one run of nphoton photons, from the batch variance
    for key in par:
        se = np.sqrt(par[key].var(0, ddof=1) * (1.0 / K + scale))
        z = (s[key] - par[key].mean(0)) / np.where(se > 0, se, np.inf)
        print("%-10s max|z| %.2f" % (key, np.abs(z).max()), flush=True)
        assert np.all(np.abs(z) < ZMAX), (key, z)
    mask = _hot_mask(par_hot)
    se = np.sqrt(par_hot[:, mask].var(0, ddof=1) * (1.0 / K + scale))
    z = (f[mask] - par_hot[:, mask].mean(0)) / se
    print("hot voxels: mean z %.3f, mean z^2 %.3f" % (z.mean(), (z * z).mean()), flush=True)
    assert (z * z).mean() < 1.6
    return s


def test_benchmark_size_run_matches_parity(batches):
    """the benchmark itself: 1e8 photons on one resident wave of 1024-lane
    workgroups (~390k photons per workgroup box)"""
    par, par_hot, _, _ = batches
    n = 100000000
    out = mcxcl_amd.run(workloads.sphere_cube_workload(nphoton=n, mathmode=0))
    assert out["stat"]["energytot"] == float(n)
    s = _single_vs_batches(out, n, par, par_hot)
    # the reference's cubesph60b absorbed fraction (speedcontest/mcxcontest:104)
    assert abs(s["absorbed"][0] - 0.2702) < 0.0005, s["absorbed"]


def test_heavily_loaded_deposit_box(batches, monkeypatch):
    """1e7 photons on 8 workgroups: 1.25M photons per float32 box, where cells
    would pass 2^20 without the move to the float64 grid"""
    par, par_hot, _, _ = batches
    n = 10000000
    monkeypatch.setenv("MCXHIP_MAXWG", "8")
    out = mcxcl_amd.run(workloads.sphere_cube_workload(nphoton=n, mathmode=0, seed=99991))
    assert out["stat"]["energytot"] == float(n)
    _single_vs_batches(out, n, par, par_hot)

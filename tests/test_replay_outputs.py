"""RF and time-weighted replay outputs and photon sharing between patterns, on
the oracle.

Replay outputs (outputtype r, s, t, b; src/mcx_core.cl:2257-2263 phase at
launch, :2568-2623 scattering-site deposits, :2845-2858 path deposits): a
replayed photon i retraces its recorded path, so each output summed over the
grid is a closed form of the forward run's records (detector id, scattering
counts n_i and partial path lengths L_ij), with w_i = exp(-sum_j mua_j L_ij)
and tof_i = sum_j L_ij n_j / c0:
  r  (RF)     -sum_i w_i L_i e^{i omega tof_i}   (real half, then imaginary half)
  s  (RFmus)   sum_i w_i n_i e^{i omega tof_i}
  t  (WLTOF)   sum_i w_i L_i tof_i
  b  (WPTOF)   sum_i w_i n_i tof_i
A phase of the wrong photon, a sign, a missing pathlen or a swapped real and
imaginary part fails these.

Photon sharing (pattern sources with srcnum > 1, src/mcx_core.cl:1678-1727,
:2902-2914): the photon launches with unit weight and deposits into one slot
per pattern, scaled by that pattern's pixel value.  When no pattern pixel is
zero no photon is relaunched, so the run draws the same random numbers as a
single-pattern run of each pattern alone, whose photons start with the pixel
value as weight: slot k equals that run's field up to float rounding."""

import math

import numpy as np
import pytest

import oracle
from mcxcl_amd import workloads

R_C0 = 1.0 / 299792458000.0   # s/mm
MUA = np.array([0.005, 0.002])
N_REF = np.array([1.37, 1.0])
OMEGA = 2 * math.pi * 100e6
NT = 4096


@pytest.fixture(scope="module")
def forward():
    base = workloads.sphere_cube_workload(nphoton=200000, issaveseed=1, savedetflag="DSP", tstep=5e-9)
    out = oracle.run(dict(base, autopilot=0, nthread=NT, nblocksize=64), nthread=NT)
    detp = out["detp"]
    assert detp.shape[1] > 500
    nscat = detp[1:3].view(np.uint32).astype(np.float64).sum(0)
    L = detp[3:5].astype(np.float64)
    rec = {
        "w": np.exp(-(MUA[:, None] * L).sum(0)),
        "L": L.sum(0),
        "n": nscat,
        "tof": (L * N_REF[:, None]).sum(0) * R_C0,
        "det": detp[0].astype(np.int64) & 0xFFFF,
    }
    return base, out, rec


def _replay(base, fwd, outputtype, **extra):
    cfg = dict(base, seed=fwd["seeds"], detphotons=fwd["detp"], outputtype=outputtype, autopilot=0, nthread=NT,
               nblocksize=64, rawfield=1, savedetflag="DSP")
    cfg.update(extra)
    return oracle.run(cfg, nthread=NT)


def _halves(out):
    raw = out["flux64"].ravel(order="F")
    h = raw.size // 2
    return raw[:h], raw[h:]


def test_rf_replay_sums(forward):
    base, fwd, r = forward
    out = _replay(base, fwd, "rf", omega=OMEGA, replaydet=0)
    re, im = _halves(out)
    ph = OMEGA * r["tof"]
    np.testing.assert_allclose(re.sum(), -(r["w"] * r["L"] * np.cos(ph)).sum(), rtol=2e-5)
    np.testing.assert_allclose(im.sum(), -(r["w"] * r["L"] * np.sin(ph)).sum(), rtol=2e-5)
    assert out["flux"].shape[-1] == 2


def test_rfmus_replay_sums(forward):
    base, fwd, r = forward
    re, im = _halves(_replay(base, fwd, "rfmus", omega=OMEGA, replaydet=0))
    ph = OMEGA * r["tof"]
    np.testing.assert_allclose(re.sum(), (r["w"] * r["n"] * np.cos(ph)).sum(), rtol=2e-5)
    np.testing.assert_allclose(im.sum(), (r["w"] * r["n"] * np.sin(ph)).sum(), rtol=2e-5)


@pytest.mark.parametrize("outputtype,key", [("wltof", "L"), ("wptof", "n")])
def test_time_weighted_replay_sums(forward, outputtype, key):
    base, fwd, r = forward
    out = _replay(base, fwd, outputtype, replaydet=0)
    np.testing.assert_allclose(out["flux64"].sum(), (r["w"] * r[key] * r["tof"]).sum(), rtol=2e-5)


def test_rf_at_zero_frequency_is_minus_the_jacobian(forward):
    base, fwd, _ = forward
    re, im = _halves(_replay(base, fwd, "rf", omega=0.0, replaydet=0))
    jac = _replay(base, fwd, "jacobian", replaydet=0)["flux64"].ravel(order="F")
    # each deposit is exactly -(w L); the float64 sums differ only in the order
    # the oracle's OpenMP workers add them
    np.testing.assert_allclose(re, -jac, rtol=1e-12, atol=0)
    assert not im.any()


def test_rf_normalisation_per_detector(forward):
    """replaydet = -1: one slice per detector in each half, both halves scaled by
    unitinmm / (summed weights of that detector's photons)
    (src/mcx_host.cpp:1397-1433)"""
    base, fwd, r = forward
    out = _replay(base, fwd, "rf", omega=OMEGA, replaydet=-1)
    re, im = _halves(out)
    flux = out["flux"]
    assert flux.shape == (60, 60, 60, 1, 4, 2)
    for d in range(1, 5):
        sw = np.float32(r["w"][r["det"] == d].sum())
        for half, part in ((re, 0), (im, 1)):
            sl = half.reshape(60, 60, 60, 4, order="F")[..., d - 1]
            np.testing.assert_allclose(flux[..., 0, d - 1, part], (sl / sw).astype(np.float32), rtol=2e-5,
                                       atol=1e-9 * np.abs(sl).max() / sw)


# ---------------- photon sharing between patterns ----------------

def _pattern_case(srcnum_patterns):
    """a 4 x 3 pixel pattern source over a 40 x 30 mm area on the top face"""
    cfg = workloads.homogeneous_cube_workload(
        nphoton=40000, srctype="pattern", srcpos=[10, 15, 0], srcparam1=[40, 0, 0, 4], srcparam2=[0, 30, 0, 3],
        autopilot=0, nthread=1024, nblocksize=64, rawfield=1)
    pats = np.stack(srcnum_patterns)                      # [srcnum, 4, 3]
    if len(srcnum_patterns) > 1:
        cfg.update(srcnum=len(srcnum_patterns), srcpattern=pats)
    else:
        cfg.update(srcpattern=pats[0])
    return cfg


def test_pattern_photon_sharing_slots_equal_single_pattern_runs():
    rng = np.random.default_rng(11)
    pa = rng.uniform(0.2, 1.0, (4, 3)).astype(np.float32)
    pb = rng.uniform(0.2, 1.0, (4, 3)).astype(np.float32)
    both = oracle.run(_pattern_case([pa, pb]), nthread=1024)
    raw = both["flux64"].ravel(order="F")
    assert both["stat"]["energytot"] == 40000.0           # unit-weight photons
    for k, pat in enumerate((pa, pb)):
        one = oracle.run(_pattern_case([pat]), nthread=1024)
        ref = one["flux64"].ravel(order="F")
        # equal up to float rounding: (w0 - w) / mua cancels in short voxel segments, so
        # per voxel the bar is absolute, 1e-5 of the field's maximum; the totals agree to 1e-6
        np.testing.assert_allclose(raw[k::2], ref, rtol=0, atol=1e-5 * ref.max())
        assert abs(raw[k::2].sum() / ref.sum() - 1) < 1e-6
        # normalisation: pattern k scaled by npixel / sum(pattern k) x the unit-photon scale
        # (src/mcx_host.cpp:1436-1447), the single run by its launched weight: the same
        # fluence up to the sample mean of the pixel weights
        fk = both["flux"].reshape(2, 60, 60, 60, order="F")[k]
        assert abs(fk.sum() / one["flux"].sum() - 1) < 0.02


def test_pattern_photon_sharing_reflectance_slots():
    """issaveref = 1 with patterns sharing photons: the escaping weight goes to
    each pattern's interleaved slot scaled by that pattern's weight
    (src/mcx_core.cl:1525-1540), so slot k is again the single-pattern run"""
    rng = np.random.default_rng(12)
    pa = rng.uniform(0.2, 1.0, (4, 3)).astype(np.float32)
    pb = rng.uniform(0.2, 1.0, (4, 3)).astype(np.float32)

    def case(pats):
        cfg = _pattern_case(pats)
        vol = np.ones((60, 60, 60), np.uint8)
        vol[:, :, 0] = 0                                   # a void layer over the tissue
        cfg.update(vol=vol, issaveref=1)
        return cfg

    both = oracle.run(case([pa, pb]), nthread=1024)
    raw = both["flux64"].ravel(order="F")
    for k, pat in enumerate((pa, pb)):
        one = oracle.run(case([pat]), nthread=1024)
        ref = one["flux64"].ravel(order="F")
        dref = ref.reshape(60, 60, 60, order="F")[:, :, 0]
        assert (dref < 0).sum() > 100                      # the reflectance tally is there
        np.testing.assert_allclose(raw[k::2], ref, rtol=0, atol=1e-5 * np.abs(ref).max())
        got = raw[k::2].reshape(60, 60, 60, order="F")[:, :, 0]
        np.testing.assert_allclose(got.sum(), dref.sum(), rtol=1e-6)

"""Several sources in one run (the reference's extrasrclen/srcid path:
src/mcx_core.cl:1602-1619, per-source field slices src/mcx_host.cpp:474, :679,
normalisation :1449-1451, source index in detid src/mcx_core.cl:887-888).

srcid 0 pools the sources: each photon draws its source uniformly.  srcid -1
does the same and deposits into one field slice per source.  srcid k >= 1
launches every photon from source k and draws nothing for it.

The CPU tests pin the oracle by what those rules fix: srcid k is the
one-source run at source k, the slices of srcid -1 add up to the pooled
field, the per-source normalisation carries the factor (number of sources),
and the source index packed into detid matches the slice a detected photon
deposited in.  The GPU tests compare the device with the oracle."""

import numpy as np
import pytest

import mcxcl_amd
import oracle
from helpers import sorted_records
from mcxcl_amd import workloads

SRC = [[29, 29, 0, 1], [12, 29, 0, 1], [45, 20, 0, 1]]
DIRS = [[0, 0, 1, 0]] * 3
NT = 2048


def _cfg(srcid, nphoton=20000, **over):
    over.setdefault("tstep", 1e-9)
    cfg = workloads.homogeneous_cube_workload(nphoton=nphoton, srcpos=SRC, srcdir=DIRS, srcid=srcid, **over)
    return cfg


def _oracle(cfg):
    return oracle.run(dict(cfg, autopilot=0, nthread=NT, nblocksize=64, rawfield=1), nthread=NT)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_fixed_source_is_the_one_source_run(k):
    multi = _oracle(_cfg(k))
    one = _oracle(workloads.homogeneous_cube_workload(nphoton=20000, srcpos=SRC[k - 1], srcdir=DIRS[k - 1],
                                                      tstep=1e-9))
    assert multi["stat"]["energytot"] == one["stat"]["energytot"]
    assert multi["stat"]["energyabs"] == one["stat"]["energyabs"]
    assert np.array_equal(multi["flux64"], one["flux64"])
    assert np.array_equal(multi["flux"], one["flux"])


def test_source_slices_add_up_to_the_pooled_field():
    pooled = _oracle(_cfg(0))
    split = _oracle(_cfg(-1))
    # [Nx, Ny, Nz, Ngate, 1 detector slot, Nsource] as the reference binding shapes it
    # (src/pmcxcl.cpp:1330-1346)
    assert pooled["flux"].ndim == 4 and split["flux"].shape == pooled["flux"].shape + (1, 3)
    assert split["stat"]["energytot"] == pooled["stat"]["energytot"]
    # the raw accumulator is [Nx, Ny, Nz, Ngate x Nsource], gate fastest
    f64 = split["flux64"].reshape(pooled["flux64"].shape + (3,), order="F")
    np.testing.assert_allclose(f64.sum(-1), pooled["flux64"], rtol=1e-12, atol=1e-300)
    # every source got photons, each holding most of the energy near its own entry voxel
    for s, (x, y, _, _) in enumerate(SRC):
        assert f64[x, y, 0:3, :, s].sum() > 0.5 * pooled["flux64"][x, y, 0:3].sum()
    # per-source normalisation: the pooled scale times the number of sources
    ratio_pooled = pooled["flux"] / np.where(pooled["flux64"] > 0, pooled["flux64"], np.inf)
    sflux = split["flux"].reshape(f64.shape, order="F")
    ratio_split = sflux / np.where(f64 > 0, f64, np.inf)
    sp = np.median(ratio_pooled[pooled["flux64"] > 0])
    ss = np.median(ratio_split[f64 > 0])
    assert ss == pytest.approx(3 * sp, rel=1e-6)


def test_detid_carries_the_source_index():
    det = [[29, 19, 0, 2], [12, 39, 0, 2], [45, 30, 0, 2]]
    cfg = _cfg(-1, nphoton=50000, detpos=det, issavedet=1, savedetflag="DP")
    out = _oracle(cfg)
    word = out["detp"][0].astype(np.int64)
    src, detid = word >> 16, word & 0xFFFF
    assert set(np.unique(src)) <= {1, 2, 3} and len(np.unique(src)) == 3
    assert set(np.unique(detid)) <= {1, 2, 3}
    # a photon from source s reaches the detector beside it far more often
    for s in (1, 2, 3):
        assert np.mean(detid[src == s] == s) > 0.5


@pytest.mark.gpu
@pytest.mark.parametrize("srcid", [0, -1, 2])
def test_multisource_parity(srcid):
    det = [[29, 19, 0, 2], [12, 39, 0, 2], [45, 30, 0, 2]]
    cfg = _cfg(srcid, nphoton=40000, detpos=det, issavedet=1, savedetflag="DSP", issaveseed=1)
    cfg = dict(cfg, autopilot=0, nthread=NT, nblocksize=64, mathmode=1, rawfield=1)
    gpu = mcxcl_amd.run(cfg)
    ref = oracle.run(cfg, nthread=NT)
    assert gpu["stat"]["energytot"] == ref["stat"]["energytot"]
    assert gpu["stat"]["energyabs"] == ref["stat"]["energyabs"]
    assert gpu["flux"].shape == ref["flux"].shape
    np.testing.assert_allclose(gpu["flux64"], ref["flux64"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(gpu["flux"], ref["flux"], rtol=1e-6, atol=0)
    assert gpu["detp"].shape == ref["detp"].shape and gpu["detp"].shape[1] > 50
    assert np.array_equal(sorted_records(gpu["detp"]), sorted_records(ref["detp"]))


def _replay_srcdet(outputtype, replaydet, fwd, base):
    return _oracle(dict(base, seed=fwd["seeds"], detphotons=fwd["detp"], outputtype=outputtype,
                        replaydet=replaydet))


@pytest.fixture(scope="module")
def multisource_forward():
    det = [[29, 19, 0, 2], [12, 39, 0, 2], [45, 30, 0, 2]]
    base = _cfg(-1, nphoton=60000, detpos=det, issavedet=1, savedetflag="DP", issaveseed=1, tstep=5e-9)
    return base, _oracle(base)


def test_replay_slices_normalised_per_source_and_detector(multisource_forward):
    """srcid -1 with replaydet -1: one Jacobian slice per (source, detector),
    each scaled by unitinmm over the replay weights of its own photons.  (The
    reference scales only the first source's slices, src/mcx_host.cpp:1398-1422.)"""
    base, fwd = multisource_forward
    word = fwd["detp"][0].astype(np.int64)
    src, det = word >> 16, word & 0xFFFF
    L = fwd["detp"][1].astype(np.float64)
    w = np.exp(-0.005 * L)
    r = _replay_srcdet("jacobian", -1, fwd, base)
    f64 = r["flux64"].reshape((60, 60, 60, 1, 3, 3), order="F")
    flux = r["flux"].reshape(f64.shape, order="F")
    for s in range(1, 4):
        for d in range(1, 4):
            sel = (src == s) & (det == d)
            sl64 = f64[..., d - 1, s - 1]
            if not sel.any():
                assert sl64.sum() == 0
                continue
            np.testing.assert_allclose(sl64.sum(), (w[sel] * L[sel]).sum(), rtol=1e-6)
            np.testing.assert_allclose(flux[..., d - 1, s - 1], (sl64 / w[sel].sum()).astype(np.float32),
                                       rtol=2e-6, atol=1e-12)


def test_replay_wp_per_detector_slices(multisource_forward):
    """WP deposits with replaydet -1: slice d equals the replaydet = d run, and the
    slices add up to the replaydet = 0 field.  The gate is clamped before the
    detector and source offsets are added; the reference clamps the sum
    (src/mcx_core.cl:2591), which with one gate sends every detector's deposits
    into the first detector's slice."""
    base, fwd = multisource_forward
    allr = _replay_srcdet("wp", -1, fwd, base)["flux64"].reshape((60, 60, 60, 1, 3, 3), order="F")
    pooled = _replay_srcdet("wp", 0, fwd, base)["flux64"].reshape((60, 60, 60, 1, 1, 3), order="F")
    np.testing.assert_allclose(allr.sum(axis=4), pooled[..., 0, :], rtol=1e-12, atol=1e-300)
    for d in (1, 3):
        one = _replay_srcdet("wp", d, fwd, base)["flux64"].reshape((60, 60, 60, 1, 1, 3), order="F")
        np.testing.assert_allclose(allr[..., d - 1, :], one[..., 0, :], rtol=1e-12, atol=1e-300)
        assert one.sum() > 0


@pytest.mark.gpu
@pytest.mark.parametrize("outputtype", ["jacobian", "wp"])
def test_multisource_replay_parity(multisource_forward, outputtype):
    base, fwd = multisource_forward
    cfg = dict(base, seed=fwd["seeds"], detphotons=fwd["detp"], outputtype=outputtype, replaydet=-1,
               autopilot=0, nthread=NT, nblocksize=64, mathmode=1, rawfield=1)
    gpu = mcxcl_amd.run(cfg)
    ref = oracle.run(cfg, nthread=NT)
    np.testing.assert_allclose(gpu["flux64"], ref["flux64"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(gpu["flux"], ref["flux"], rtol=1e-6, atol=1e-12)


@pytest.mark.gpu
def test_srcid_minus_two_device_parity():
    """srcid -2: the detectors launch as disk sources beside the user's source, one
    fluence per source, device equal to the oracle (bit-reproducible mode)"""
    cfg = workloads.homogeneous_cube_workload(nphoton=30000, srcpos=[29.5, 29.5, 0], tstep=1e-9)
    cfg.update(outputtype="fluence", srcid=-2, detpos=[[20, 29.5, 0, 2], [39, 29.5, 0, 2]],
               detdir=[[0, 0, 1, 0], [0, 0, 1, 0]], issavedet=0, autopilot=0, nthread=2048, nblocksize=64,
               mathmode=1, rawfield=1)
    gpu, ref = mcxcl_amd.run(cfg), oracle.run(cfg, nthread=2048)
    assert gpu["stat"]["energytot"] == ref["stat"]["energytot"]
    assert gpu["stat"]["energyabs"] == ref["stat"]["energyabs"]
    np.testing.assert_allclose(gpu["flux64"], ref["flux64"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(gpu["flux"], ref["flux"], rtol=1e-6, atol=0)
    assert gpu["flux"].shape[-1] == 3

"""Adjoint (phi_src . phi_det) outputs, outputtype adjoint* (src/mcx_host.cpp:1467-1642,
src/mcx_core.cl:2159-2184 and :3313-3512).  Detectors are appended as disk sources,
one run launches every source (srcid -1), and per voxel and (source, detector)
pair the outputs are
  jmua  = -Vvox     * sum_t phi_s * sum_t phi_d
  jd    = -unitinmm * grad sum_t phi_s . grad sum_t phi_d
  jmus / jmusp = jd / (3 (1-g) mus^2) / (3 (1-g)^2 mus^2) of the voxel's label.

The reference holds no golden vectors for these (its adjoint demos compare with
an external FEM solver), so the oracle's post-processing is pinned by closed
forms: exact differences of quadratic fields, numpy.gradient (second-order
edges) in float64, the per-label factors, and the disk launch of the detector
sources by a nearly ballistic beam."""

import numpy as np
import pytest

import mcxcl_amd
import oracle
from mcxcl_amd import pmcxcl, workloads

NT = 4096


def _case(outputtype="adjoint_mua_d", nphoton=40000, **over):
    cfg = workloads.homogeneous_cube_workload(nphoton=nphoton, srcpos=[29.5, 29.5, 0], tstep=5e-9)
    cfg.update(outputtype=outputtype, detpos=[[20, 29.5, 0, 2], [39, 29.5, 0, 2]],
               detdir=[[0, 0, 1, 0], [0, 0, 1, 0]], issavedet=0)
    cfg.update(over)
    return cfg


def test_detectors_become_disk_sources():
    with pmcxcl.prepare(_case()) as pc:
        assert pc.c.extrasrclen == 2 and pc.c.srcid == -1 and pc.nsrc == 3
    with pmcxcl.prepare(_case(outputtype="flux")) as pc:
        assert pc.c.extrasrclen == 0


def test_standalone_validation_appends_the_detectors():
    """a stand-alone run (parentid mpStandalone) gets the same append from
    mcx_validatecfg itself (src/mcx_utils.c:1902-1925)"""
    import ctypes
    from mcxcl_amd import _native
    c = pmcxcl.Config()
    _native.lib.mcx_initcfg(ctypes.byref(c))
    try:
        pmcxcl._parse_config(_case(), c)
        c.parentid = 0
        assert _native.lib.mcx_try_validatecfg(ctypes.byref(c), None, None, 0) == 0
        assert c.extrasrclen == 2 and c.srcid == -1
        sd = np.ctypeslib.as_array((ctypes.c_float * 32).from_address(c.srcdata)).reshape(2, 4, 4)
        assert np.allclose(sd[:, 2, 0], [2, 2]) and np.allclose(sd[:, 1, :3], [[0, 0, 1], [0, 0, 1]])
        assert np.allclose(sd[:, 0, 3], [1, 1])
    finally:
        _native.lib.mcx_clearcfg(ctypes.byref(c))


def test_difference_stencil_is_exact_on_quadratics():
    """second-order central and one-sided differences differentiate a quadratic
    exactly; a 2-voxel axis falls back to the first difference"""
    x = np.arange(7, dtype=np.float32)
    q = (3 * x * x - 5 * x + 2).astype(np.float32)
    c3 = np.broadcast_to(q[:, None, None], (7, 3, 2)).astype(np.float32)
    g = oracle._fd_grad(c3, 0)
    assert np.array_equal(g[:, 0, 0], (6 * x - 5).astype(np.float32))
    assert np.array_equal(oracle._fd_grad(c3, 1), np.zeros_like(c3))          # constant along y
    two = np.array([[[1.0], [4.0]]], np.float32)                             # a 2-voxel y axis
    assert np.array_equal(oracle._fd_grad(two, 1)[0, :, 0], np.array([3.0, 3.0], np.float32))
    assert np.array_equal(oracle._fd_grad(np.ones((1, 2, 2), np.float32), 0), np.zeros((1, 2, 2), np.float32))


@pytest.mark.parametrize("outputtype,keys,tstep", [("adjoint_mua_d", ("jmua", "jd"), 5e-9),
                                                   ("adjoint_mua_musp", ("jmua", "jmusp"), 5e-9),
                                                   ("adjoint_mus", ("jmus",), 5e-9),
                                                   ("adjoint_mua_d", ("jmua", "jd"), 1e-9)])
def test_oracle_outputs_match_closed_forms(outputtype, keys, tstep):
    """with one gate and with five (the products use the gate sums)"""
    cfg = _case(outputtype, prop=[[0, 0, 1, 1], [0.005, 1.0, 0.8, 1.37]], tstep=tstep)
    out = oracle.run(cfg, nthread=NT)
    assert set(keys) <= set(out) and out[keys[0]].shape == (60, 60, 60, 2)
    ng = int(round(5e-9 / tstep))
    phi = out["flux"].reshape(60, 60, 60, ng, 3, order="F").astype(np.float64).sum(3)   # [x, y, z, source]
    assert (phi[..., 1:] > 0).any()
    grads = [np.gradient(phi[..., k], edge_order=2) for k in range(3)]
    musp_factor = 1.0 / (3.0 * (1 - 0.8) ** 2 * 1.0 ** 2)
    mus_factor = 1.0 / (3.0 * (1 - 0.8) * 1.0 ** 2)
    for d in range(2):
        j_mua = -1.0 * phi[..., 0] * phi[..., 1 + d]
        j_d = -1.0 * sum(gs * gd for gs, gd in zip(grads[0], grads[1 + d]))
        want = {"jmua": j_mua, "jd": j_d, "jmusp": j_d * musp_factor, "jmus": j_d * mus_factor}
        for k in keys:
            got = out[k][..., d].astype(np.float64)
            scale = np.abs(want[k]).max()
            assert scale > 0
            np.testing.assert_allclose(got, want[k], rtol=1e-4, atol=1e-5 * scale)


def test_detector_sources_launch_from_a_disk():
    """a nearly ballistic medium (mus 1e-4/mm): the detector source's fluence fills
    a cylinder of the detector radius along its direction, evenly across it"""
    cfg = _case("adjoint", nphoton=60000, prop=[[0, 0, 1, 1], [0.01, 1e-4, 0.0, 1.0]],
                detpos=[[30, 30, 0, 6]], detdir=[[0, 0, 1, 0]], srcpos=[5.5, 5.5, 0])
    out = oracle.run(cfg, nthread=NT)
    det = out["flux"].reshape(60, 60, 60, -1, order="F")[..., 1].astype(np.float64).sum(2)   # [x, y]
    xc = np.arange(60) + 0.5
    r = np.hypot(xc[:, None] - 30, xc[None, :] - 30)
    # outside the disk only the tails of the ~0.6% of photons that scatter over 60 mm
    assert det[r > 6 + 0.75].sum() <= 1e-2 * det.sum()
    inner = det[r < 5]
    assert inner.min() > 0 and inner.std() / inner.mean() < 0.25
    xbar = (det * xc[:, None]).sum() / det.sum()
    ybar = (det * xc[None, :]).sum() / det.sum()
    assert abs(xbar - 30) < 0.2 and abs(ybar - 30) < 0.2


ADJOINT_TYPES = ["adjoint", "adjoint_dcoeff", "adjoint_mus", "adjoint_musp", "adjoint_mua_d", "adjoint_mua_musp"]


@pytest.mark.gpu
@pytest.mark.parametrize("outputtype,tstep", [(t, 5e-9) for t in ADJOINT_TYPES] + [("adjoint_mua_musp", 1e-9)])
def test_adjoint_outputs_device_parity(outputtype, tstep):
    """device against the oracle in bit-reproducible mode: the per-source field
    as for any run, the adjoint products within float32 rounding (one gate, and
    five for one type)"""
    cfg = dict(_case(outputtype, nphoton=20000, prop=[[0, 0, 1, 1], [0.005, 1.0, 0.8, 1.37]], tstep=tstep),
               autopilot=0, nthread=2048, nblocksize=64, mathmode=1, rawfield=1)
    gpu, ref = mcxcl_amd.run(cfg), oracle.run(cfg, nthread=2048)
    assert gpu["stat"]["energytot"] == ref["stat"]["energytot"]
    np.testing.assert_allclose(gpu["flux64"], ref["flux64"], rtol=1e-9, atol=1e-12)
    keys = [k for k in ("jmua", "jd", "jmus", "jmusp") if k in ref]
    assert keys and [k for k in ("jmua", "jd", "jmus", "jmusp") if k in gpu] == keys
    for k in keys:
        scale = np.abs(ref[k]).max()
        assert scale > 0
        np.testing.assert_allclose(gpu[k], ref[k], rtol=1e-5, atol=1e-6 * scale)


def test_srcid_minus_two_stacks_the_detector_sources():
    """srcid -2 (src/mcx_utils.c:1902-1925, src/mcx_core.cl:2159-2160): the detectors
    are appended as disk sources as for the adjoint outputs, srcid stays -2, and the
    run returns one fluence per source with no adjoint products: bit for bit the
    per-source fluence an adjoint-output run computes from the same photons"""
    fl = _case(outputtype="fluence", srcid=-2)
    with pmcxcl.prepare(fl) as pc:
        assert pc.c.extrasrclen == 2 and pc.c.srcid == -2 and pc.nsrc == 3
    a = oracle.run(dict(fl, autopilot=0, nthread=NT, nblocksize=64, rawfield=1), nthread=NT)
    b = oracle.run(dict(_case(), autopilot=0, nthread=NT, nblocksize=64, rawfield=1), nthread=NT)
    assert "jmua" not in a and "jmua" in b
    assert np.array_equal(a["flux64"], b["flux64"])
    assert a["flux"].shape == (60, 60, 60, 1, 1, 3)      # [x, y, z, gate, 1, source] (src/pmcxcl.cpp:1339-1345)
    np.testing.assert_array_equal(a["flux"], b["flux"])

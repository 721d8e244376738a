"""RF forward simulation (omega > 0 without replay; src/mcx_core.cl:2427-2430,
:2750-2760, :2833-2839, :3033-3040) on the CPU oracle.

Photon weights are complex, w <- w exp(-(mua + i omega n / c0) ds); a voxel
exit deposits (w0 - w) / (mua + i omega n / c0), the real part in the first
half of the grid and the imaginary part in the second; the binding returns
the normalised fluence as complex64.  Pins:
  * closed form: a pencil beam through a non-scattering slab deposits
    N e^{-a z} (1 - e^{-a}) / a in voxel z, a = mua + i omega n / c0 (the sign
    of the phase, the complex division and the layout of the two halves);
  * the omega -> 0 limit: the real part is the ordinary flux run's grid to
    float rounding and the imaginary part vanishes, on a scattering volume
    with reflection;
  * the complex output is the grid times the flux normaliser."""

import numpy as np
import pytest

import mcxcl_amd
import oracle
from mcxcl_amd import workloads

NT = 64
R_C0 = 3.335640951981520e-12


def _run(cfg, **kw):
    return oracle.run(dict(cfg, autopilot=0, nthread=NT, nblocksize=64, rawfield=1, **kw), nthread=NT)


def test_non_scattering_slab_closed_form():
    cfg = workloads.homogeneous_cube_workload(nphoton=1000, tstep=5e-9, isreflect=0)
    cfg["prop"] = [[0, 0, 1, 1], [0.01, 0.0, 0.0, 1.37]]
    om = 2 * np.pi * 1e9
    out = _run(cfg, omega=om)
    f = out["flux64"]
    assert f.shape == (60, 60, 60, 2)
    z = np.arange(60)
    a = 0.01 + 1j * om * 1.37 * R_C0
    want = 1000 * np.exp(-a * z) * (1 - np.exp(-a)) / a
    got = f[29, 29, :, 0] + 1j * f[29, 29, :, 1]
    np.testing.assert_allclose(got, want, rtol=2e-5)
    # nothing anywhere else
    assert np.count_nonzero(f) == 2 * 60
    # the phase turns by omega n L / c0 over the slab: about 1.72 rad
    assert np.angle(got[-1]) == pytest.approx(-(om * 1.37 * R_C0 * 59.5), abs=0.02)


def test_zero_frequency_limit():
    cfg = workloads.sphere_cube_workload(nphoton=20000, tstep=1e-9)
    base = _run(cfg)
    rf = _run(cfg, omega=1e-3)
    assert rf["stat"]["energyabs"] == pytest.approx(base["stat"]["energyabs"], rel=1e-6)
    re, im = rf["flux64"][..., :5], rf["flux64"][..., 5:]
    np.testing.assert_allclose(re, base["flux64"], rtol=1e-5, atol=1e-9 * np.abs(base["flux64"]).max())
    assert np.abs(im).max() <= 1e-9 * np.abs(re).max()


def test_complex_output_is_the_normalised_grid():
    cfg = workloads.homogeneous_cube_workload(nphoton=5000, tstep=5e-9)
    om = 2 * np.pi * 100e6
    out = _run(cfg, omega=om)
    flux = out["flux"]
    assert flux.dtype == np.complex64 and flux.shape == (60, 60, 60, 1)
    g = out["flux64"]
    scale = 1.0 / (out["stat"]["energytot"] * 5e-9)     # unitinmm / (energytot Vvox tstep)
    np.testing.assert_allclose(flux.real, (g[..., 0:1] * scale).astype(np.float32), rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(flux.imag, (g[..., 1:2] * scale).astype(np.float32), rtol=1e-5, atol=1e-6)
    assert (flux.imag <= 1e-3).all() and flux.imag.min() < -1e-4     # the phase lags


def test_rf_forward_pattern_sources_stay_real():
    """a pattern source's deposit has no imaginary term (src/mcx_core.cl:2902-2914):
    the imaginary half stays zero, and with patterns sharing photons slot k is the
    single-pattern run of pattern k"""
    rng = np.random.default_rng(4)
    pa = rng.uniform(0.2, 1.0, (4, 3)).astype(np.float32)
    pb = rng.uniform(0.2, 1.0, (4, 3)).astype(np.float32)

    def case(pats):
        cfg = workloads.homogeneous_cube_workload(nphoton=20000, srctype="pattern", srcpos=[10, 15, 0],
                                                  srcparam1=[40, 0, 0, 4], srcparam2=[0, 30, 0, 3], tstep=5e-9)
        if len(pats) > 1:
            cfg.update(srcnum=len(pats), srcpattern=np.stack(pats))
        else:
            cfg.update(srcpattern=pats[0])
        return _run(cfg, omega=2 * np.pi * 100e6)

    both = case([pa, pb])
    raw = both["flux64"].ravel(order="F")
    half = raw.size // 2
    assert not raw[half:].any() and np.abs(raw[:half]).max() > 0
    for k, pat in enumerate((pa, pb)):
        one = case([pat])
        ref = one["flux64"].ravel(order="F")
        assert not ref[ref.size // 2:].any()
        np.testing.assert_allclose(raw[:half][k::2], ref[:ref.size // 2], rtol=0,
                                   atol=1e-5 * np.abs(ref).max())


def _adjoint_case(outputtype, omega, nphoton=40000):
    cfg = workloads.homogeneous_cube_workload(nphoton=nphoton, srcpos=[29.5, 29.5, 0], tstep=5e-9)
    cfg.update(outputtype=outputtype, detpos=[[20, 29.5, 0, 2], [39, 29.5, 0, 2]],
               detdir=[[0, 0, 1, 0], [0, 0, 1, 0]], issavedet=0, omega=omega)
    return cfg


def test_pooled_sources_slices_add_to_the_pooled_field():
    """srcid -1 with complex weights: the per-source complex fields add up to the
    srcid 0 field of the same photons (the source is drawn per photon either way)"""
    base = workloads.homogeneous_cube_workload(nphoton=20000, tstep=5e-9, srcpos=[[29, 29, 0], [15, 40, 0]],
                                               srcdir=[[0, 0, 1], [0, 0, 1]], omega=2 * np.pi * 200e6)
    pooled, split = _run(dict(base, srcid=0)), _run(dict(base, srcid=-1))
    assert split["flux64"].shape == (60, 60, 60, 4)          # [re s1, re s2, im s1, im s2]
    re = split["flux64"][..., 0] + split["flux64"][..., 1]
    im = split["flux64"][..., 2] + split["flux64"][..., 3]
    np.testing.assert_allclose(re, pooled["flux64"][..., 0], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(im, pooled["flux64"][..., 1], rtol=1e-9, atol=1e-12)
    assert split["flux"].dtype == np.complex64 and split["flux"].shape == (60, 60, 60, 1, 1, 2)


def test_complex_adjoint_products():
    """RF adjoint outputs (src/mcx_core.cl:3390-3511, src/mcx_host.cpp:1467-1642): the
    plain (not conjugated) complex products of the normalised complex fluences,
    against a complex128 computation from the returned per-source field"""
    out = _run(_adjoint_case("adjoint_mua_d", 2 * np.pi * 300e6))
    flux = out["flux"][:, :, :, 0, 0, :]                      # [x, y, z, source], complex, one gate
    assert flux.shape == (60, 60, 60, 3) and np.abs(flux.imag).max() > 0
    phi = flux.astype(np.complex128)
    jmua = -1.0 * phi[..., :1, None] * phi[..., None, 1:]     # Vvox = 1
    jmua = jmua.reshape(60, 60, 60, 2)
    assert out["jmua"].dtype == np.complex64
    np.testing.assert_allclose(out["jmua"], jmua, rtol=1e-4, atol=1e-6 * np.abs(jmua).max())
    grad = lambda f: np.stack(np.gradient(f, edge_order=2), axis=-1)
    gs = grad(phi[..., 0])
    for d in range(2):
        jd = -1.0 * (gs * grad(phi[..., 1 + d])).sum(-1)
        np.testing.assert_allclose(out["jd"][..., d], jd, rtol=2e-3, atol=1e-4 * np.abs(jd).max())


def test_complex_adjoint_zero_frequency_limit():
    cw = _run(_adjoint_case("adjoint", 0.0))
    rf = _run(_adjoint_case("adjoint", 1e-3))
    np.testing.assert_allclose(rf["jmua"].real, cw["jmua"], rtol=1e-4, atol=1e-7 * np.abs(cw["jmua"]).max())
    assert np.abs(rf["jmua"].imag).max() <= 1e-8 * np.abs(cw["jmua"]).max()

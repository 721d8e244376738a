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


def test_rf_forward_restrictions():
    cfg = workloads.sphere_cube_workload(nphoton=100)
    with pytest.raises(RuntimeError, match="RF forward"):
        # pooled sources (srcid -1) stack one field per source: not built for complex fields
        mcxcl_amd.run(dict(cfg, omega=1e8, srcpos=[[29, 29, 0], [20, 20, 0]], srcdir=[[0, 0, 1], [0, 0, 1]],
                           srcid=-1))

"""Device RF forward simulation (omega > 0 without replay, src/mcx_core.cl:
2750-2760, :2833-2839): complex photon weights on the full general path,
against the oracle with the same seeds and thread map (bit-reproducible mode):
energies and records bit-exact, both halves of the float64 grid within 1e-9
relative (atomic summation order only), the complex output within 1e-6.  A
fast-mode run agrees with the bit-reproducible one within Monte Carlo error,
and the non-scattering slab matches its closed form on the device too."""

import numpy as np
import pytest

import mcxcl_amd
import oracle
from helpers import sorted_records
from mcxcl_amd import workloads

pytestmark = pytest.mark.gpu

NT = 2048
OMEGA = 2 * np.pi * 150e6
R_C0 = 3.335640951981520e-12


def test_rf_forward_parity():
    cfg = workloads.sphere_cube_workload(nphoton=50000, tstep=1e-9, savedetflag="DPX", issaveseed=1,
                                         minenergy=0.01)
    cfg = dict(cfg, autopilot=0, nthread=NT, nblocksize=64, mathmode=1, rawfield=1, omega=OMEGA)
    gpu, ref = mcxcl_amd.run(cfg), oracle.run(cfg, nthread=NT)
    assert gpu["stat"]["energytot"] == ref["stat"]["energytot"]
    assert gpu["stat"]["energyabs"] == ref["stat"]["energyabs"]
    assert np.abs(ref["flux64"][..., 5:]).max() > 0
    np.testing.assert_allclose(gpu["flux64"], ref["flux64"], rtol=1e-9, atol=1e-12 * np.abs(ref["flux64"]).max())
    assert gpu["flux"].dtype == np.complex64
    np.testing.assert_allclose(gpu["flux"], ref["flux"], rtol=1e-6, atol=1e-9 * np.abs(ref["flux"]).max())
    assert np.array_equal(sorted_records(gpu["detp"]), sorted_records(ref["detp"]))


def test_rf_forward_fast_mode():
    cfg = workloads.sphere_cube_workload(nphoton=400000, tstep=1e-9, omega=OMEGA)
    n = cfg["nphoton"]
    fast = mcxcl_amd.run(dict(cfg, mathmode=0))
    slow = mcxcl_amd.run(dict(cfg, mathmode=1))
    af, ap = (o["stat"]["energyabs"] / n for o in (fast, slow))
    assert abs(af - ap) <= 6 * np.sqrt(2 * ap * (1 - ap) / n), (af, ap)
    for part in (np.real, np.imag):
        sf, sp = float(part(fast["flux"]).sum()), float(part(slow["flux"]).sum())
        assert abs(sf / sp - 1) < 0.01, (sf, sp)


def test_rf_forward_slab_closed_form():
    cfg = workloads.homogeneous_cube_workload(nphoton=1000, tstep=5e-9, isreflect=0)
    cfg["prop"] = [[0, 0, 1, 1], [0.01, 0.0, 0.0, 1.37]]
    om = 2 * np.pi * 1e9
    out = mcxcl_amd.run(dict(cfg, mathmode=0, rawfield=1, omega=om))
    f = out["flux64"]
    z = np.arange(60)
    a = 0.01 + 1j * om * 1.37 * R_C0
    want = 1000 * np.exp(-a * z) * (1 - np.exp(-a)) / a
    np.testing.assert_allclose(f[29, 29, :, 0] + 1j * f[29, 29, :, 1], want, rtol=2e-4)


@pytest.mark.parametrize("outputtype", ["adjoint_mua_d", "adjoint_mua_musp", "adjoint_mus"])
def test_complex_adjoint_device_parity(outputtype):
    """RF forward with the detectors as disk sources: the complex per-source field
    and the complex adjoint products, device against the oracle (bit-reproducible
    mode): the grid within 1e-9, the products within float32 rounding"""
    cfg = workloads.homogeneous_cube_workload(nphoton=20000, srcpos=[29.5, 29.5, 0], tstep=5e-9,
                                              prop=[[0, 0, 1, 1], [0.005, 1.0, 0.8, 1.37]])
    cfg.update(outputtype=outputtype, detpos=[[20, 29.5, 0, 2], [39, 29.5, 0, 2]],
               detdir=[[0, 0, 1, 0], [0, 0, 1, 0]], issavedet=0, omega=OMEGA, autopilot=0, nthread=NT,
               nblocksize=64, mathmode=1, rawfield=1)
    gpu, ref = mcxcl_amd.run(cfg), oracle.run(cfg, nthread=NT)
    assert gpu["stat"]["energytot"] == ref["stat"]["energytot"]
    np.testing.assert_allclose(gpu["flux64"], ref["flux64"], rtol=1e-9, atol=1e-12 * np.abs(ref["flux64"]).max())
    keys = [k for k in ("jmua", "jd", "jmus", "jmusp") if k in ref]
    assert keys and [k for k in ("jmua", "jd", "jmus", "jmusp") if k in gpu] == keys
    for k in keys:
        assert gpu[k].dtype == np.complex64
        scale = np.abs(ref[k]).max()
        assert scale > 0 and np.abs(ref[k].imag).max() > 0
        np.testing.assert_allclose(gpu[k], ref[k], rtol=1e-5, atol=1e-6 * scale)


def test_rf_forward_shared_patterns_parity():
    """patterns sharing photons with complex weights: real deposits per pattern
    slot, no imaginary term (src/mcx_core.cl:2902-2914)"""
    pats = np.random.default_rng(6).uniform(0.2, 1.0, (2, 4, 3)).astype(np.float32)
    cfg = workloads.sphere_cube_workload(nphoton=30000, tstep=1e-9, srctype="pattern", srcpos=[10, 15, 0],
                                         srcparam1=[40, 0, 0, 4], srcparam2=[0, 30, 0, 3], srcnum=2,
                                         srcpattern=pats)
    cfg = dict(cfg, autopilot=0, nthread=NT, nblocksize=64, mathmode=1, rawfield=1, omega=OMEGA)
    gpu, ref = mcxcl_amd.run(cfg), oracle.run(cfg, nthread=NT)
    assert gpu["stat"]["energyabs"] == ref["stat"]["energyabs"]
    raw = ref["flux64"].ravel(order="F")
    assert np.abs(raw).max() > 0 and not raw[raw.size // 2:].any()
    np.testing.assert_allclose(gpu["flux64"], ref["flux64"], rtol=1e-9, atol=1e-12 * np.abs(raw).max())
    np.testing.assert_allclose(gpu["flux"], ref["flux"], rtol=1e-6, atol=1e-9 * np.abs(ref["flux"]).max())

"""Pins of the oracle's output types and time gating against closed forms.

A non-scattering column (1 x 1 x 40 voxels, mus so small that no photon
scatters, mua = 0.1/mm, n = 1.5 inside and out, no reflection) gives every
photon the same deterministic history: it enters voxel k with weight
w_k = exp(-0.1 k) and leaves it at t = (k + 1) n / c0.  So the deposits of the
reference's output types (src/mcx_core.cl:2833-2861) are known exactly:
  energy   (w0 - w)          N (w_k - w_k+1)
  flux     (w0 - w) / mua    N (w_k - w_k+1) / mua
  length   w0 L              N w_k
each in gate floor(t_k / tstep).  A deposit using the exit weight where the
entry weight belongs, a path length in the wrong unit, or a time of flight
without the refractive index fails these.

The DCS replay deposit ((1 - cos theta) w per scattering event, :2568-2623)
summed over the grid equals sum_i w_i sum_j M_ij, the momentum transfers the
forward run recorded for the same photons (savedetflag M)."""

import math

import numpy as np
import pytest

import oracle
from mcxcl_amd import workloads

R_C0 = 1.0 / 299792458000.0   # s/mm
N = 5000


def _column(outputtype):
    cfg = workloads.homogeneous_cube_workload(
        nphoton=N, vol=np.ones((1, 1, 40), np.uint8), srcpos=[0.5, 0.5, 0], isreflect=0,
        prop=[[0, 0, 1, 1.5], [0.1, 1e-7, 0.0, 1.5]], tstart=0.0, tend=2e-10, tstep=1e-11,
        outputtype=outputtype, autopilot=0, nthread=128, nblocksize=64, rawfield=1)
    out = oracle.run(cfg, nthread=128)
    assert out["counters"]["nscat"] == 0
    return out["flux64"].reshape(40, -1)


@pytest.mark.parametrize("outputtype", ["energy", "flux", "length"])
def test_column_deposits_and_gates(outputtype):
    field = _column(outputtype)
    ngate = field.shape[1]
    assert ngate == 20
    k = np.arange(40)
    w = np.exp(-0.1 * np.arange(41))
    expect = {"energy": N * (w[:-1] - w[1:]), "flux": N * (w[:-1] - w[1:]) / 0.1, "length": N * w[:-1]}[outputtype]
    gate = np.floor((k + 1) * 1.5 * R_C0 / 1e-11).astype(int)
    inside = gate < ngate
    assert inside.sum() > 30
    want = np.zeros_like(field)
    want[k[inside], gate[inside]] = expect[inside]
    np.testing.assert_allclose(field, want, rtol=2e-5, atol=0)


def test_dcs_replay_sums_to_recorded_momentum_transfer():
    base = workloads.sphere_cube_workload(nphoton=100000, tstep=5e-9, savedetflag="DSPM", issaveseed=1)
    nt = 4096
    fwd = oracle.run(dict(base, autopilot=0, nthread=nt, nblocksize=64), nthread=nt)
    detp = fwd["detp"]
    assert detp.shape[1] > 200
    L = detp[3:5].astype(np.float64)
    mom = detp[5:7].astype(np.float64)
    w = np.exp(-(np.array([0.005, 0.002])[:, None] * L).sum(0))
    rep = dict(base, seed=fwd["seeds"], detphotons=detp, outputtype="wm", replaydet=0, autopilot=0, nthread=nt,
               nblocksize=64, rawfield=1)
    dcs = oracle.run(rep, nthread=nt)["flux64"]
    np.testing.assert_allclose(dcs.sum(), (w * mom.sum(0)).sum(), rtol=2e-5)
    # per medium: the sphere (label 2) and the rest
    vol = base["vol"]
    np.testing.assert_allclose(dcs[..., 0][vol == 2].sum(), (w * mom[1]).sum(), rtol=2e-5)
    np.testing.assert_allclose(dcs[..., 0][vol == 1].sum(), (w * mom[0]).sum(), rtol=2e-5)

"""Batched launch (include/mcx_host.h mcx_batch_*): many independent
simulations in one persistent kernel.  The reference has no counterpart (its
host runs one Config per call), so each instance of a batch is checked
against the same Config run alone through mcx_run_simulation, whose own
parity with the oracle the other device tests establish.

Exact checks: every instance launches exactly its photon count (each launch
adds the source weight 1 to the launched-energy tally) and a replayed HIP
graph keeps that bookkeeping on every replay.  Statistical checks: absorbed
fractions and fluence sums agree with the single runs within Monte Carlo
error."""

import numpy as np
import pytest

from mcxcl_amd import pmcxcl, workloads


def _small_batch(n=12, dim=32, seed=5, photons=(1e4, 6e4)):
    return workloads.batched_workload(ninst=n, dim=dim, seed=seed, photons=photons)


@pytest.mark.parametrize("bad,msg", [
    ({"mathmode": 1}, "fast arithmetic"),
    ({"issavedet": 1, "detpos": [[16, 16, 0, 2]]}, "detected photons"),
    ({"debuglevel": "R"}, "RNG dump"),
])
def test_batch_rejects_unsupported_instances(bad, msg):
    cfgs = _small_batch(n=3)
    cfgs[1] = dict(cfgs[1], **bad)
    with pytest.raises(RuntimeError, match=msg):
        pmcxcl.run_batch(cfgs)


def test_batch_rejects_bad_shapes():
    with pytest.raises(RuntimeError, match="blocksize"):
        pmcxcl.run_batch(_small_batch(n=2), blocksize=96)
    with pytest.raises(RuntimeError, match="at least one"):
        pmcxcl.run_batch([])


def test_batched_workload_is_seeded_and_spans_the_ranges():
    a = workloads.batched_workload(ninst=40, dim=32, seed=9)
    b = workloads.batched_workload(ninst=40, dim=32, seed=9)
    assert all(np.array_equal(x["vol"], y["vol"]) and x["prop"] == y["prop"] for x, y in zip(a, b))
    props = np.array([p for c in a for p in c["prop"][1:]])
    assert props[:, 0].min() >= 0.001 and props[:, 0].max() <= 0.05
    assert props[:, 1].min() >= 0.5 and props[:, 1].max() <= 20
    assert props[:, 2].min() >= 0.7 and props[:, 2].max() <= 0.95
    n = np.array([c["nphoton"] for c in a])
    assert n.min() >= 1e4 and n.max() <= 1e6
    assert {c["srctype"] for c in a} == {"pencil", "isotropic", "disk"}
    assert all(1 <= (c["vol"] == 2).sum() < c["vol"].size for c in a)


def _check_against_single(outs, cfgs, nsig=6.0):
    for out, cfg in zip(outs, cfgs):
        n = cfg["nphoton"]
        st = out["stat"]
        assert st["energytot"] == n, (cfg["session"], st["energytot"], n)
        single = pmcxcl.run(cfg)
        a_b = st["energyabs"] / st["energytot"]
        a_s = single["stat"]["energyabs"] / single["stat"]["energytot"]
        sig = np.sqrt(max(a_s * (1 - a_s), 1e-4) / n)
        assert abs(a_b - a_s) <= nsig * np.sqrt(2) * sig + 1e-4, (cfg["session"], a_b, a_s, sig)
        fb, fs = float(out["flux"].sum()), float(single["flux"].sum())
        assert abs(fb - fs) <= 0.05 * abs(fs) + 1e-6, (cfg["session"], fb, fs)


@pytest.mark.gpu
def test_batch_matches_single_runs():
    cfgs = _small_batch()
    outs, info = pmcxcl.run_batch(cfgs)
    print("schedule", info, flush=True)
    assert info["nblock"] > 0 and info["probe"]
    _check_against_single(outs, cfgs)


@pytest.mark.gpu
def test_batch_fixed_schedule_without_graph():
    cfgs = _small_batch(n=6, seed=11)
    outs, info = pmcxcl.run_batch(cfgs, blocksize=128, probe=False, usegraph=False)
    assert info["blocksize"] == 128 and not info["probe"] and info["kernelms"] > 0
    _check_against_single(outs, cfgs)


@pytest.mark.gpu
def test_batch_wide_workgroups():
    """one 1024-lane workgroup per CU, each with a deposit box sized from the
    CU's LDS, gives the same instances as the single runs"""
    cfgs = _small_batch(n=6, seed=17)
    outs, info = pmcxcl.run_batch(cfgs, blocksize=1024)
    assert info["blocksize"] == 1024 and info["nblock"] > 0
    _check_against_single(outs, cfgs)


@pytest.mark.gpu
def test_graph_replays_keep_bookkeeping():
    """replay the captured step several times, with eager work and
    synchronisations in between; every replay must launch every photon and
    agree with the eager (uncaptured) batch"""
    cfgs = _small_batch(n=8, seed=13)
    eager, _ = pmcxcl.run_batch(cfgs, usegraph=False)
    with pmcxcl.Batch(cfgs, usegraph=True) as b:
        for rep in range(3):
            b.run(1)
            outs = b.collect()
            pmcxcl.run(cfgs[rep])           # eager work between replays
            for out, ref, cfg in zip(outs, eager, cfgs):
                n = cfg["nphoton"]
                assert out["stat"]["energytot"] == n
                a, e = (out["stat"]["energyabs"] / n, ref["stat"]["energyabs"] / n)
                assert abs(a - e) <= 6 * np.sqrt(2 * max(e * (1 - e), 1e-4) / n) + 1e-4, (rep, cfg["session"], a, e)
                assert np.isfinite(out["flux"]).all()


@pytest.mark.gpu
def test_batch_of_identical_instances_averages_to_the_single_run():
    base = workloads.homogeneous_cube_workload(nphoton=200000, srcpos=[29.5, 29.5, 29.5], srctype="isotropic")
    cfgs = [dict(base, seed=100 + i) for i in range(8)]
    outs, _ = pmcxcl.run_batch(cfgs)
    single = pmcxcl.run(dict(base, nphoton=1600000))
    mean = np.mean([o["flux"] for o in outs], axis=0)
    ref = single["flux"]
    # the fluence near the source (well sampled) agrees within a few percent
    core = ref[25:35, 25:35, 25:35]
    np.testing.assert_allclose(mean[25:35, 25:35, 25:35].sum(), core.sum(), rtol=0.02)

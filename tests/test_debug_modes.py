"""Move-only trajectories (debug flag T, src/mcx_core.cl:3287-3291, host
src/mcx_utils.c:1552-1555) and fill-max detection (issavedet = 3,
src/mcx_core.cl:921-923, :1484-1489) on the CPU oracle.

Pins:
  * move-only records are the trajectory run's (flag M) records thread by
    thread: identical when no photon enters an in-grid background voxel, a
    per-thread prefix when one does (the thread's run ends there, without its
    energy write-back); no fluence and no detected photons are recorded;
  * fill-max mode on one worker keeps the first maxdetphoton detections of the
    ordinary run, record for record, and launches fewer photons.

The reference's Python binding takes the debug letters R, M and P only; T comes
from the command line or a JSON input, so these tests widen the binding's
letters for the run."""

import numpy as np
import pytest

import oracle
from helpers import sorted_records
from mcxcl_amd import pmcxcl, workloads

NT = 16


@pytest.fixture(autouse=True)
def _cli_letters(monkeypatch):
    monkeypatch.setattr(pmcxcl, "DEBUGFLAGS", "RMPT")


def _slab(void_layer):
    N = 10
    vol = np.ones((20, 20, N), np.uint8)
    if void_layer:
        vol[:, :, N - 1] = 0
    return dict(nphoton=400, vol=vol, prop=[[0, 0, 1, 1.0], [0.01, 1.0, 0.9, 1.0]], srcpos=[10.5, 10.5, 0],
                srcdir=[0, 0, 1], tstart=0, tend=5e-9, tstep=5e-9, isreflect=1, issrcfrom0=1, seed=5,
                autopilot=0, nthread=NT, nblocksize=NT)


def _by_thread(tr, nphoton):
    """records grouped by the thread that wrote them, in the order written
    (one worker: threads run one after another)"""
    ids = tr[0].view(np.uint32)
    tp, odd = nphoton // NT, nphoton % NT
    first = [t * tp + min(t, odd) + 1 for t in range(NT + 1)]
    # launch/scatter ids can differ from termination ids past oddphoton; with odd = 0 they agree
    assert odd == 0
    return [tr[:, (ids >= first[t]) & (ids < first[t + 1])] for t in range(NT)]


def test_move_only_equals_trajectories_without_background_entry():
    cfg = _slab(False)
    m = oracle.run(dict(cfg, debuglevel="M"), nthread=NT, nomp=1)
    t = oracle.run(dict(cfg, debuglevel="T"), nthread=NT, nomp=1)
    assert np.array_equal(t["traj"].view(np.uint32), m["traj"].view(np.uint32))
    assert not t["flux64"].any() and t["stat"]["detected"] == 0
    assert t["stat"]["energytot"] == m["stat"]["energytot"] == 400.0   # threads that finish write back


def test_move_only_ends_each_thread_at_its_first_background_entry():
    cfg = _slab(True)
    m = oracle.run(dict(cfg, debuglevel="M"), nthread=NT, nomp=1)
    t = oracle.run(dict(cfg, debuglevel="T"), nthread=NT, nomp=1)
    assert t["traj"].shape[1] < m["traj"].shape[1]
    ended = 0
    for a, b in zip(_by_thread(t["traj"], 400), _by_thread(m["traj"], 400)):
        n = a.shape[1]
        assert np.array_equal(a.view(np.uint32), b[:, :n].view(np.uint32))    # a prefix
        ended += n < b.shape[1]
    assert ended == NT                                     # every thread met the background layer
    assert t["stat"]["energytot"] == 0.0                   # no thread wrote its energies back


def test_fill_max_keeps_the_first_detections():
    cfg = workloads.sphere_cube_workload(nphoton=20000, tstep=1e-9)
    cfg = dict(cfg, autopilot=0, nthread=256, nblocksize=64)
    full = oracle.run(cfg, nthread=256, nomp=1)
    fill = oracle.run(dict(cfg, issavedet=3, maxdetphoton=50), nthread=256, nomp=1)
    assert full["stat"]["detected"] > 50
    assert fill["stat"]["detected"] == 50
    assert np.array_equal(fill["detp"].view(np.uint32), full["detp"][:, :50].view(np.uint32))
    assert fill["stat"]["energytot"] < full["stat"]["energytot"]
    assert np.array_equal(sorted_records(fill["detp"]), sorted_records(full["detp"][:, :50]))

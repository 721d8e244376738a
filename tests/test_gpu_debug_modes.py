"""Device runs of move-only trajectories (debug flag T, src/mcx_core.cl:3287-3291)
and fill-max detection (issavedet = 3, src/mcx_core.cl:921-923, :1484-1489).

Move-only: every lane's run depends on that lane alone, so in bit-reproducible
mode the records match the oracle's as a set, and so do the energies of the
lanes that finished.  Fill-max: lanes stop as the shared count fills, so which
photons are detected depends on timing; each record is one of the ordinary
run's records (bit-reproducible mode, same thread map), the count is the
buffer size and fewer photons are launched."""

import numpy as np
import pytest

import mcxcl_amd
import oracle
from helpers import sorted_records
from mcxcl_amd import pmcxcl, workloads

pytestmark = pytest.mark.gpu

NT = 512


@pytest.fixture(autouse=True)
def _cli_letters(monkeypatch):
    monkeypatch.setattr(pmcxcl, "DEBUGFLAGS", "RMPT")


def _rows(a):
    return {r.tobytes() for r in np.ascontiguousarray(np.asarray(a, np.float32).T)}


def test_move_only_parity():
    vol = np.ones((20, 20, 10), np.uint8)
    vol[:, :, 9] = 0
    cfg = dict(nphoton=20000, vol=vol, prop=[[0, 0, 1, 1.0], [0.01, 1.0, 0.9, 1.0]], srcpos=[10.5, 10.5, 0],
               srcdir=[0, 0, 1], tstart=0, tend=5e-9, tstep=5e-9, isreflect=1, issrcfrom0=1, seed=5,
               autopilot=0, nthread=NT, nblocksize=64, mathmode=1, debuglevel="T")
    gpu, ref = mcxcl_amd.run(cfg), oracle.run(cfg, nthread=NT)
    assert ref["traj"].shape[1] > NT
    assert gpu["traj"].shape == ref["traj"].shape
    assert np.array_equal(sorted_records(gpu["traj"]), sorted_records(ref["traj"]))
    assert gpu["stat"]["energytot"] == ref["stat"]["energytot"]


@pytest.mark.parametrize("mathmode", [1, 0])
def test_fill_max(mathmode):
    # fast mode runs one 1024-lane workgroup per CU whatever nthread says: enough
    # photons that lanes are still launching when the buffer fills
    n = 200000 if mathmode == 1 else 5000000
    cfg = workloads.sphere_cube_workload(nphoton=n, tstep=1e-9)
    cfg = dict(cfg, autopilot=0, nthread=4096, nblocksize=64, mathmode=mathmode)
    fill = mcxcl_amd.run(dict(cfg, issavedet=3, maxdetphoton=300))
    assert fill["stat"]["detected"] == 300 and fill["detp"].shape[1] == 300
    assert fill["stat"]["energytot"] < n
    if mathmode == 1:
        full = mcxcl_amd.run(cfg)
        assert _rows(fill["detp"]) <= _rows(full["detp"])

"""Device trajectory records (debug flag M, src/mcx_core.cl:929-944): the
records of launch, scattering and termination events, bit-exact against the
oracle as a set (the device appends them in the order lanes reach the
counter) in bit-reproducible mode, on the sphere workload with detectors and
reflection.  Fast arithmetic keeps the fixed thread map when recording
trajectories, so its termination records also carry photon ids 1..nphoton."""

import numpy as np
import pytest

import mcxcl_amd
import oracle
from mcxcl_amd import workloads

pytestmark = pytest.mark.gpu

NT = 2048


def _sorted(tr):
    keys = np.ascontiguousarray(tr.T).view(np.uint32).reshape(-1, 6)
    return keys[np.lexsort(keys.T[::-1])]


def test_trajectory_parity():
    cfg = workloads.sphere_cube_workload(nphoton=3000, tstep=1e-9, savedetflag="DPX", issaveseed=1)
    cfg = dict(cfg, autopilot=0, nthread=NT, nblocksize=64, mathmode=1, debuglevel="M", rawfield=1)
    gpu, ref = mcxcl_amd.run(cfg), oracle.run(cfg, nthread=NT)
    c = ref["counters"]
    assert ref["traj"].shape == (6, int(2 * c["nlaunch"] + c["nscat"]))
    assert gpu["traj"].shape == ref["traj"].shape
    assert np.array_equal(_sorted(gpu["traj"]), _sorted(ref["traj"]))
    # the fluence is unchanged by recording
    np.testing.assert_allclose(gpu["flux64"], ref["flux64"], rtol=1e-9, atol=1e-12)


def test_trajectory_fast_mode_and_cap():
    cfg = workloads.sphere_cube_workload(nphoton=20000, tstep=1e-9)
    out = mcxcl_amd.run(dict(cfg, mathmode=0, debuglevel="M"))
    tr = out["traj"]
    ids = tr[0].view(np.uint32)
    # launch records: at the source with unit weight, one per photon
    launch = (tr[4] == 1.0) & np.all(tr[1:4] == np.float32(cfg["srcpos"])[:, None], axis=0)
    assert launch.sum() == 20000
    assert ids.max() <= 20000 and ids.min() >= 1
    assert out["stat"]["energytot"] == 20000
    capped = mcxcl_amd.run(dict(cfg, mathmode=0, debuglevel="M", maxjumpdebug=1000))
    assert capped["traj"].shape == (6, 1000)

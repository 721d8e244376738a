"""Trajectory records (debug flag M, src/mcx_core.cl:929-944, :1497-1503,
:2243-2249, :2625-2633) on the CPU oracle.

Each record is {photon id as uint32 bits, x, y, z, w, source id}, written when a
photon is launched, at every scattering event and when it terminates.  Pins:
  * count: records = 2 x launches + scattering events (the oracle's own event
    counters), and maxjumpdebug caps what is stored;
  * termination records carry the photon ids 1..nphoton once each (the photon
    count of the thread map, min(idx, oddphoton));
  * launch and scattering records carry the reference's id min(idx, (idx <
    oddphoton) idx): equal to the termination id below oddphoton, idx*threadphoton
    + f.w past it;
  * geometry: a pencil beam launches at the source with unit weight; along one
    photon the weight never grows (absorption only, no roulette), consecutive
    scattering sites lie inside the volume, and in a non-scattering medium each
    photon has exactly its launch and its exit, on the far face."""

import numpy as np

import oracle
from mcxcl_amd import workloads

NT = 64


def _run(cfg, **kw):
    return oracle.run(dict(cfg, autopilot=0, nthread=NT, nblocksize=64, debuglevel="M", **kw), nthread=NT)


def _split(tr):
    ids = tr[0].view(np.uint32)
    return ids, tr[1:4], tr[4], tr[5]


def test_record_count_and_cap():
    cfg = workloads.homogeneous_cube_workload(nphoton=100, tstep=5e-9)
    out = _run(cfg)
    c = out["counters"]
    assert out["traj"].shape == (6, int(2 * c["nlaunch"] + c["nscat"]))
    capped = _run(cfg, maxjumpdebug=50)
    assert capped["traj"].shape == (6, 50)


def test_photon_ids():
    cfg = workloads.homogeneous_cube_workload(nphoton=100, tstep=5e-9)
    # one worker: threads run in order, each thread's photons one after another
    out = oracle.run(dict(cfg, autopilot=0, nthread=NT, nblocksize=64, debuglevel="M"), nthread=NT, nomp=1)
    ids, pos, w, src = _split(out["traj"])
    assert (src == 0).all()
    # every launch record sits at the source with unit weight
    starts = np.flatnonzero((w == 1.0) & np.all(pos == np.float32([29, 29, 0])[:, None], axis=0))
    assert len(starts) == 100
    ends = list(starts[1:] - 1) + [len(w) - 1]
    # 100 photons on 64 threads: threadphoton 1, oddphoton 36
    tp, odd, k = 1, 36, 0
    for idx in range(NT):
        for fw in range(1, tp + (idx < odd) + 1):
            a, b = starts[k], ends[k]
            # launch and scattering records: the reference's min(idx, (idx < oddphoton) idx)
            assert (ids[a:b] == idx * tp + (idx if idx < odd else 0) + fw).all()
            # termination record: the photon's number in the thread map, 1..nphoton
            assert ids[b] == idx * tp + min(idx, odd) + fw == k + 1
            k += 1


def test_weights_never_grow_along_a_photon():
    cfg = workloads.homogeneous_cube_workload(nphoton=200, tstep=5e-9)
    ids, pos, w, _ = _split(oracle.run(dict(cfg, autopilot=0, nthread=1, nblocksize=1, debuglevel="M"),
                                       nthread=1)["traj"])
    # one thread: records in time order; a photon's records are contiguous
    starts = np.flatnonzero(w == 1.0)
    assert len(starts) == 200
    for a, b in zip(starts, list(starts[1:]) + [len(w)]):
        seg = w[a:b]
        assert (np.diff(seg) <= 0).all()
        inner = pos[:, a + 1:b - 1]     # scattering sites
        assert (inner >= 0).all() and (inner <= 60).all()


def test_non_scattering_photons_have_launch_and_exit_only():
    cfg = workloads.homogeneous_cube_workload(nphoton=50, tstep=5e-9, isreflect=0)
    cfg["prop"] = [[0, 0, 1, 1], [0.01, 0.0, 0.0, 1.0]]
    out = _run(cfg)
    ids, pos, w, _ = _split(out["traj"])
    assert out["traj"].shape[1] == 100
    launch = w == 1.0
    assert launch.sum() == 50
    exitp = pos[:, ~launch]
    # straight down the z axis through 60 voxels of mua 0.01
    np.testing.assert_allclose(exitp[2], 60.0, atol=1e-4)
    np.testing.assert_allclose(w[~launch], np.exp(-0.6), rtol=1e-5)

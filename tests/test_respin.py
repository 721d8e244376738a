"""respin (-r/--repeat, src/mcx_host.cpp:1089-1104, :1280-1332): |respin|
repetitions of nphoton/|respin| photons, each on the next nthread*4 rand() words
of the same srand(seed) stream; the field, the energy tallies and the detected
photons add up over the repetitions.

The reference's host sums the repetitions' fields with a double count (each
repetition's read-back of the device field already holds the earlier ones,
src/mcx_host.cpp:1250-1297 with the field cleared only between time windows,
:1312); this build and the oracle sum each repetition once.

CPU pins: the RNG dump of the second repetition equals the xorshift128+ streams
(an independent Python statement) of the second block of libc rand() words; the
first repetition is exactly a one-repetition run of nphoton/respin photons; the
launched energy counts every photon once.  GPU: device equals oracle."""

import numpy as np
import pytest

import mcxcl_amd
import oracle
from helpers import sorted_records, xorshift128p_floats
from mcxcl_amd import workloads

NT = 256
SEED = 1648335518


def _small(**over):
    vol = np.ones((20, 20, 20), np.uint8)
    return workloads.homogeneous_cube_workload(vol=vol, srcpos=[10, 10, 0], autopilot=0, nthread=NT, nblocksize=64,
                                               mathmode=1, **over)


def test_second_repetition_draws_the_next_seed_words():
    out = oracle.run(_small(nphoton=1000, debuglevel="R", respin=2), nthread=NT)
    field = out["flux"].ravel(order="F")
    words = oracle.glibc_rand(SEED, 8 * NT)[4 * NT:]       # libc: the second block of nthread*4 words
    expect = np.zeros_like(field)
    for t in range(NT):
        stream = xorshift128p_floats(words[4 * t:4 * t + 4], (field.size - t + NT - 1) // NT)
        expect[t::NT] = stream
    assert np.array_equal(field.view(np.uint32), expect.view(np.uint32))


def test_first_repetition_is_the_shorter_run():
    """respin 2 = a run of nphoton/2 photons on the first seeds plus a second, nonnegative part"""
    two = oracle.run(_small(nphoton=20001, respin=2, rawfield=1), nthread=NT)
    one = oracle.run(_small(nphoton=10000, rawfield=1), nthread=NT)
    assert two["stat"]["energytot"] == 20000.0       # 2 x floor(20001 / 2)
    assert one["stat"]["energytot"] == 10000.0
    rest = two["flux64"] - one["flux64"]
    assert rest.min() > -1e-9 * two["flux64"].max()
    # the second part carries about half of the absorbed energy
    ea = two["stat"]["energyabs"] - one["stat"]["energyabs"]
    assert abs(ea / one["stat"]["energyabs"] - 1) < 0.05


def test_negative_respin_splits_the_same_total():
    out = oracle.run(_small(nphoton=30000, respin=-3), nthread=NT)
    assert out["stat"]["energytot"] == 30000.0


@pytest.mark.gpu
def test_respin_parity():
    cfg = workloads.sphere_cube_workload(nphoton=90001, respin=3, issaveseed=1, savedetflag="DP", autopilot=0,
                                         nthread=2048, nblocksize=64, mathmode=1, rawfield=1)
    gpu = mcxcl_amd.run(cfg)
    ref = oracle.run(cfg, nthread=2048)
    assert gpu["stat"]["energytot"] == ref["stat"]["energytot"] == 90000.0
    assert gpu["stat"]["energyabs"] == ref["stat"]["energyabs"]
    np.testing.assert_allclose(gpu["flux64"], ref["flux64"], rtol=1e-9, atol=1e-12)
    assert gpu["detp"].shape == ref["detp"].shape and gpu["detp"].shape[1] > 50
    assert np.array_equal(sorted_records(gpu["detp"]), sorted_records(ref["detp"]))


@pytest.mark.gpu
def test_respin_rng_dump_parity():
    cfg = _small(nphoton=1000, debuglevel="R", respin=2)
    gpu = mcxcl_amd.run(cfg)["flux"]
    ref = oracle.run(cfg, nthread=NT)["flux"]
    assert np.array_equal(gpu.view(np.uint32), ref.view(np.uint32))


@pytest.mark.gpu
def test_respin_fast_mode_launches_every_photon():
    n = 2000000
    two = mcxcl_amd.run(workloads.homogeneous_cube_workload(nphoton=n, respin=2))
    one = mcxcl_amd.run(workloads.homogeneous_cube_workload(nphoton=n, seed=7))
    assert two["stat"]["energytot"] == float(n)
    a2 = two["stat"]["energyabs"] / n
    a1 = one["stat"]["energyabs"] / n
    assert abs(a2 - a1) < 2e-3, (a2, a1)

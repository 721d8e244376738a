"""The command-line tool end to end on the device: output files and replay
from a saved detected-photon file, as the reference's shell tests run it
(test/testmcx.sh:115-128)."""

import os
import re
import subprocess

import numpy as np
import pytest

from mcxcl_amd import files

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mcxcl_amd", "lib", "mcxcl")

pytestmark = pytest.mark.gpu


def _mcxcl(*args):
    print("mcxcl", " ".join(args), flush=True)
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def _count(out, word):
    return [int(float(x)) for x in re.findall(word + r"\s+([0-9.]+) photons", out)]


def test_replay_from_a_jdat_file_detects_every_replayed_photon(tmp_path):
    """run cube60 saving seeds, then replay its _detp.jdat: the replay run
    detects every photon it simulates and absorbs 30-38% (the reference's
    '-E' replay checks compare those two counts of the replay run)"""
    root = str(tmp_path)
    first = _mcxcl("--bench", "cube60", "-s", "replaytest", "-q", "1", "-S", "0", "--root", root)
    ndet = _count(first, "detected")[-1]
    assert ndet > 1000
    assert os.path.exists(os.path.join(root, "replaytest_detp.jdat"))
    again = _mcxcl("--bench", "cube60", "-E", os.path.join(root, "replaytest_detp.jdat"), "-S", "0", "--root", root)
    nrep = _count(again, "simulated")[-1]
    assert _count(again, "detected")[-1] == nrep
    # the .jdat loader drops photons whose time of flight, summed from the
    # partial paths, lands outside [tstart, tend] (src/mcx_utils.c:1455-1457):
    # at most a few that the kernel's own clock still counted in the window
    assert ndet - 5 <= nrep <= ndet
    absorbed = float(re.search(r"absorbed:\s*([0-9.]+)%", again).group(1))
    assert 30.0 <= absorbed <= 38.0


def test_replay_from_an_mch_file(tmp_path):
    root = str(tmp_path)
    first = _mcxcl("--bench", "cube60", "-s", "mchrun", "-q", "1", "-S", "0", "-F", "mc2", "--root", root)
    ndet = _count(first, "detected")[-1]
    data, heads, seeds = files.loadmch(os.path.join(root, "mchrun.mch"))
    assert data.shape[0] == ndet and seeds.shape == (ndet, 16)
    assert heads[0]["detected"] == ndet and heads[0]["totalphoton"] == 1000000
    again = _mcxcl("--bench", "cube60", "-E", os.path.join(root, "mchrun.mch"), "-S", "0", "--root", root)
    assert _count(again, "detected")[-1] == ndet


def test_fluence_formats_hold_the_same_field(tmp_path):
    """the bit-reproducible mode gives the same fluence run after run, so every
    format must hold the same numbers"""
    root = str(tmp_path)
    vols = {}
    for fmt in ("mc2", "nii", "jnii", "bnii"):
        _mcxcl("--bench", "cube60", "-n", "200000", "--mathmode", "1", "-F", fmt, "-s", "f_" + fmt, "--root", root)
    vols["mc2"] = files.loadmc2(os.path.join(root, "f_mc2.mc2"), [60, 60, 60, 1])
    vols["nii"] = files.loadnii(os.path.join(root, "f_nii.nii"))[0]
    vols["jnii"] = files.loadjnii(os.path.join(root, "f_jnii.jnii"))[0]
    vols["bnii"] = files.loadbnii(os.path.join(root, "f_bnii.bnii"))[0]
    ref = vols["mc2"]
    assert ref.sum() > 0
    for k, v in vols.items():
        assert np.array_equal(np.asarray(v).reshape(60, 60, 60, -1, order="F")[..., 0], ref[..., 0]), k

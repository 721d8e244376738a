"""Published absorbed-fraction / detection anchors of the reference test suite
(test/testmcx.sh), checked on the CPU oracle and on the gfx950 kernel."""

import numpy as np
import pytest

import mcxcl_amd
from mcxcl_amd import anchors
import oracle

ANCHORS = anchors.anchors()
# the CPU suite runs the reference's CI photon count (test/testmcx.sh is run with -n 1e5)
CPU_PHOTONS = 100000


def _quantity(out, kind):
    if kind == "absorbed":
        st = out["stat"]
        return st["energyabs"] / st["energytot"]
    return out["detp"].shape[1] if "detp" in out else 0


@pytest.mark.parametrize("name", sorted(ANCHORS))
def test_anchor_oracle(name):
    cfg, kind, lo, hi, src = ANCHORS[name]
    cfg = dict(cfg)
    if cfg.get("nphoton", 0) > CPU_PHOTONS or name in ("cube60", "cube60b"):
        cfg["nphoton"] = min(cfg.get("nphoton", 1000000), CPU_PHOTONS) if "nphoton" in cfg else CPU_PHOTONS
    q = _quantity(oracle.run(cfg, nthread=16384), kind)
    assert lo <= q < hi, "%s: %s=%r outside [%r, %r) (%s)" % (name, kind, q, lo, hi, src)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(ANCHORS))
def test_anchor_gpu(name):
    cfg, kind, lo, hi, src = ANCHORS[name]
    q = _quantity(mcxcl_amd.run(dict(cfg)), kind)
    assert lo <= q < hi, "%s: %s=%r outside [%r, %r) (%s)" % (name, kind, q, lo, hi, src)

"""Published absorbed-fraction anchors of the reference's command-line test suite.

Each entry restates one check of test/testmcx.sh as a pmcxcl-style config and the
interval of absorbed fraction (or detected count) the reference accepts.  The
configs start from the cube60 benchmark (src/mcx_bench.h:31-90) with the same
--json modifications the shell script applies.
"""

import json

import numpy as np

from . import workloads

CUBE60 = workloads.bench["cube60"]


def _cube60(**over):
    cfg = dict(CUBE60)
    cfg["vol"] = np.ones([60, 60, 60], dtype=np.uint8)
    cfg["issavedet"] = 0
    cfg.pop("detpos", None)
    cfg.update(over)
    return cfg


def anchors():
    """name -> (config, quantity, lo, hi, source)"""
    a = {}
    a["cube60"] = (_cube60(), "absorbed", 0.17, 0.18, "test/testmcx.sh:61")
    a["cube60b"] = (_cube60(isreflect=1), "absorbed", 0.27, 0.28, "test/testmcx.sh:65")
    a["cube60_bc_aarraa"] = (_cube60(isreflect=0, bc="aarraa"), "absorbed", 0.27, 0.28, "test/testmcx.sh:73")
    a["cube60_cyclic"] = (_cube60(bc="cccccc", nphoton=1000), "absorbed", 0.99, 1.0, "test/testmcx.sh:77")
    a["cube60planar"] = (_cube60(isreflect=1, srctype="planar", srcpos=[10, 10, -10], srcparam1=[40, 0, 0, 0],
                                 srcparam2=[0, 40, 0, 0]), "absorbed", 0.25, 0.26, "test/testmcx.sh:85")
    a["isotropic"] = (_cube60(srctype="isotropic", srcpos=[29, 29, 29]), "absorbed", 0.88, 0.89,
                      "test/testmcx.sh:89")
    a["cone"] = (_cube60(prop=[[0, 0, 1, 1], [0.001, 0.001, 0, 1]], srctype="cone", srcparam1=[0.5, 0, 0, 0]),
                 "absorbed", 0.06, 0.07, "test/testmcx.sh:93")
    a["fourier"] = (_cube60(isreflect=1, srctype="fourier", srcpos=[10, 10, -10], srcparam1=[40, 0, 0, 2],
                            srcparam2=[0, 40, 0, 0]), "absorbed", 0.25, 0.26, "test/testmcx.sh:97")
    a["pencilarray"] = (_cube60(isreflect=1, srctype="pencilarray", srcpos=[10, 10, -10], srcparam1=[40, 0, 0, 4],
                                srcparam2=[0, 20, 0, 2]), "absorbed", 0.23, 0.24, "test/testmcx.sh:101")
    a["bc_detect"] = (_cube60(bc="______111111", issavedet=1, nphoton=10000), "detected", 9700, 10000,
                      "test/testmcx.sh:105")
    shapes2d = json.dumps({"Shapes": [{"Grid": {"Tag": 1, "Size": [1, 100, 100]}},
                                      {"Box": {"Tag": 2, "O": [0, 30, 10], "Size": [1, 40, 40]}}]})
    a["2d"] = (_cube60(isreflect=1, vol=np.ones([1, 100, 100], dtype=np.uint8), shapes=shapes2d,
                       prop=[[0, 0, 1, 1], [0.02, 0.1, 0.9, 1.37], [0.02, 10, 0.9, 6.85]], srcpos=[0, 50, 0]),
               "absorbed", 0.60, 0.70, "test/testmcx.sh:109")
    skin = json.dumps({"Shapes": [{"Grid": {"Tag": 1, "Size": [200, 200, 200]}},
                                  {"ZLayers": [[1, 20, 1], [21, 32, 4], [33, 200, 3]]},
                                  {"Cylinder": {"Tag": 2, "C0": [0, 100.5, 100.5], "C1": [200, 100.5, 100.5],
                                                "R": 20}}]})
    a["skinvessel"] = ({
        "nphoton": 100000, "seed": 1648335518, "isreflect": 1, "tstart": 0, "tend": 5e-8, "tstep": 5e-8,
        "srcpos": [100, 100, 20], "srcdir": [0, 0, 1], "srcparam1": [60, 0, 0, 0], "srctype": "disk",
        "issrcfrom0": 1, "unitinmm": 0.005, "vol": np.ones([200, 200, 200], dtype=np.uint8), "shapes": skin,
        "prop": [[1e-05, 0, 1, 1.37], [3.564e-05, 1, 1, 1.37], [23.05426549, 9.398496241, 0.9, 1.37],
                 [0.04584957865, 35.65405549, 0.9, 1.37], [1.657237447, 37.59398496, 0.9, 1.37]],
        "issavedet": 0}, "absorbed", 0.39, 0.40, "test/testmcx.sh:113")
    return a

"""Host-side configuration layer (the reference's src/mcx_utils.c and
src/mcx_shapes.c): input files, presets, shape rasterisation, detector
masking and record geometry.  CPU only."""

import ctypes
import json
import os

import numpy as np
import pytest

from mcxcl_amd import _native, pmcxcl

lib = _native.lib
QT = os.path.join(os.path.dirname(__file__), "golden", "quicktest")
DET_MASK = 0x80000000


def _arr(ptr, n, ct=ctypes.c_float):
    return np.ctypeslib.as_array((ct * n).from_address(ptr)).copy()


def _readconfig(path):
    c = _native.Config()
    lib.mcx_initcfg(ctypes.byref(c))
    cwd = os.getcwd()
    os.chdir(os.path.dirname(path))
    try:
        rc = lib.mcx_try_readconfig(os.path.basename(path).encode(), ctypes.byref(c))
    finally:
        os.chdir(cwd)
    assert rc == 0, _native.last_error()
    return c


def _summary(c):
    nvox = c.dim.x * c.dim.y * c.dim.z
    return {
        "nphoton": c.nphoton,
        "srcpos": (c.srcpos.x, c.srcpos.y, c.srcpos.z, c.srcpos.w),
        "srcdir": (c.srcdir.x, c.srcdir.y, c.srcdir.z),
        "time": (c.tstart, c.tend, c.tstep),
        "dim": (c.dim.x, c.dim.y, c.dim.z),
        "prop": _arr(c.prop, 4 * c.medianum).tolist(),
        "detpos": _arr(c.detpos, 4 * c.detnum).tolist(),
        "vol": _arr(c.vol, nvox, ctypes.c_uint),
    }


def test_quicktest_inp_and_json_describe_the_same_run():
    """example/quicktest: qtest.inp (1-based positions, mus g mua n media rows) and
    qtest.json (OriginType 1) load to the same Config"""
    a = _readconfig(os.path.join(QT, "qtest.inp"))
    b = _readconfig(os.path.join(QT, "qtest.json"))
    sa, sb = _summary(a), _summary(b)
    for key in ("nphoton", "srcpos", "srcdir", "time", "dim", "prop", "detpos"):
        assert sa[key] == sb[key], key
    assert np.array_equal(sa["vol"], sb["vol"])
    assert sa["srcpos"] == (29.0, 29.0, 0.0, 1.0)
    assert sa["prop"][4:] == pytest.approx([0.005, 1.0, 0.01, 1.0])
    # the shape file is a 60^3 grid of label 1; detector voxels carry bit 31
    labels = sa["vol"] & 0x7FFFFFFF
    assert labels.min() == 1 and labels.max() == 1
    assert (sa["vol"] & DET_MASK).any()
    # src/mcx_utils.c mcx_loadconfig reads the .inp seed only when the seed is 0;
    # mcx_initcfg sets 0x623F9A9E, so the file's seed is skipped.  JSON sets it.
    assert a.seed == 0x623F9A9E and b.seed == 29012392
    for c in (a, b):
        lib.mcx_clearcfg(ctypes.byref(c))


@pytest.mark.parametrize("name", ["cube60", "cube60b", "cube60planar", "cubesph60b", "skinvessel", "sphshells"])
def test_presets_load_and_validate(name):
    c = _native.Config()
    lib.mcx_initcfg(ctypes.byref(c))
    assert lib.mcx_try_loadbenchmark(name.encode(), ctypes.byref(c)) == 0, _native.last_error()
    assert lib.mcx_try_validatecfg(ctypes.byref(c), None, None, 0) == 0, _native.last_error()
    assert c.vol and c.medianum >= 2 and c.nphoton > 0
    if name in ("cube60", "cube60b", "cube60planar", "cubesph60b"):
        assert (c.dim.x, c.dim.y, c.dim.z) == (60, 60, 60)
        assert ord(c.isreflect) == (0 if name == "cube60" else 1)
    lib.mcx_clearcfg(ctypes.byref(c))


def _prepared_labels(shapes, dim, issrcfrom0=0):
    p = pmcxcl.prepare({"nphoton": 10, "vol": np.zeros(dim, np.uint8), "shapes": json.dumps(shapes),
                        "issrcfrom0": issrcfrom0,
                        "prop": [[0, 0, 1, 1], [0.01, 1, 0, 1],
                                                                                  [0.02, 2, 0, 1], [0.03, 3, 0, 1]],
                        "srcpos": [1, 1, 1], "srcdir": [0, 0, 1], "tstart": 0, "tend": 1e-9, "tstep": 1e-9})
    try:
        c = p.c
        assert (c.dim.x, c.dim.y, c.dim.z) == dim
        v = _arr(c.vol, dim[0] * dim[1] * dim[2], ctypes.c_uint) & 0x7FFFFFFF
        return v.reshape(dim, order="F")
    finally:
        p.close()


@pytest.mark.parametrize("issrcfrom0", [0, 1])
def test_shape_sphere_and_box_rasterisation(issrcfrom0):
    """voxel (i,j,k) belongs to a sphere when |(i,j,k)+0.5-O| <= R
    (src/mcx_shapes.c:268-280, no origin shift) and to a box when every centre
    coordinate is within [O-orig, O-orig+Size] (:396-417), orig = (1,1,1) unless
    issrcfrom0 (src/mcx_utils.c:1942-1946); later shapes overwrite earlier ones"""
    dim = (40, 36, 30)
    shapes = {"Shapes": [{"Grid": {"Tag": 1, "Size": list(dim)}},
                         {"Sphere": {"Tag": 2, "O": [20, 18, 15], "R": 9.5}},
                         {"Box": {"Tag": 3, "O": [2, 3, 4], "Size": [10, 12, 8]}}]}
    got = _prepared_labels(shapes, dim, issrcfrom0)
    i, j, k = np.meshgrid(np.arange(dim[0]), np.arange(dim[1]), np.arange(dim[2]), indexing="ij")
    want = np.ones(dim, np.uint32)
    sph = (i + 0.5 - 20) ** 2 + (j + 0.5 - 18) ** 2 + (k + 0.5 - 15) ** 2 <= 9.5 ** 2
    want[sph] = 2
    o = 0 if issrcfrom0 else 1
    x0, y0, z0 = 2 - o, 3 - o, 4 - o
    box = ((i + 0.5 >= x0) & (i + 0.5 <= x0 + 10) & (j + 0.5 >= y0) & (j + 0.5 <= y0 + 12) & (k + 0.5 >= z0) &
           (k + 0.5 <= z0 + 8))
    want[box] = 3
    assert np.array_equal(got, want)


def test_shape_layers_and_cylinder():
    dim = (20, 20, 30)
    shapes = {"Shapes": [{"Grid": {"Tag": 1, "Size": list(dim)}},
                         {"ZLayers": [[1, 5, 2], [6, 10, 3]]},
                         {"Cylinder": {"Tag": 2, "C0": [10, 10, 0], "C1": [10, 10, 30], "R": 4}}]}
    got = _prepared_labels(shapes, dim, issrcfrom0=1)
    # ZLayers rows are [start, end, tag] with 1-based inclusive z indices
    assert (got[0, 0, 0:5] == 2).all() and (got[0, 0, 5:10] == 3).all() and (got[0, 0, 10:] == 1).all()
    i, j = np.meshgrid(np.arange(dim[0]), np.arange(dim[1]), indexing="ij")
    inside = (i + 0.5 - 10) ** 2 + (j + 0.5 - 10) ** 2 <= 16
    assert (got[inside][:, 12] == 2).all()
    assert (got[~inside][:, 12] == 1).all()


def _expected_detmask(labels, dets):
    """src/mcx_utils.c:4085-4198: sample a half-voxel lattice within R+1 of each
    detector; a sampled voxel is marked when all its corners lie within R+sqrt(3)
    of the centre, its nearest corner within R+0.5, it is not background and one
    of its 26 neighbours (outside the grid counts as background) is background."""
    nx, ny, nz = labels.shape
    pad = np.zeros((nx + 2, ny + 2, nz + 2), labels.dtype)
    pad[1:-1, 1:-1, 1:-1] = labels
    mark = np.zeros(labels.shape, bool)
    corners = np.array([[a, b, c] for a in (0, 1) for b in (0, 1) for c in (0, 1)], np.float32)
    for (cx, cy, cz, r) in dets:
        steps = np.arange(-r - 1.0, r + 1.0 + 1e-6, 0.5, dtype=np.float32)
        for z in steps:
            for y in steps:
                for x in steps:
                    ix, iy, iz = x + cx, y + cy, z + cz
                    if ix < 0 or iy < 0 or iz < 0 or ix >= nx or iy >= ny or iz >= nz:
                        continue
                    if x * x + y * y + z * z > (r + 1) * (r + 1):
                        continue
                    v = np.array([int(ix), int(iy), int(iz)], np.float32)
                    d2 = (((v - [cx, cy, cz])[None, :] + corners) ** 2).sum(1)
                    if (d2 > (r + 1.7321) ** 2).any() or d2.min() >= (r + 0.5) ** 2:
                        continue
                    a, b, c = int(ix) + 1, int(iy) + 1, int(iz) + 1
                    if pad[a, b, c] and not pad[a - 1:a + 2, b - 1:b + 2, c - 1:c + 2].all():
                        mark[a - 1, b - 1, c - 1] = True
    return mark


def test_detector_mask_matches_rule():
    vol = np.ones((30, 30, 30), np.uint8)
    vol[10:20, 10:20, 12:] = 2
    vol[:, :, 25:] = 0
    dets = [[15.0, 5.0, 0.0, 2.0], [4.5, 20.0, 0.0, 1.0], [15.0, 15.0, 25.0, 3.0]]
    p = pmcxcl.prepare({"nphoton": 10, "vol": vol, "prop": [[0, 0, 1, 1], [0.01, 1, 0, 1.37], [0.02, 2, 0, 1.37]],
                        "srcpos": [15, 15, 0], "srcdir": [0, 0, 1], "tstart": 0, "tend": 1e-9, "tstep": 1e-9,
                        "detpos": dets, "issavedet": 1, "issrcfrom0": 1})
    try:
        c = p.c
        v = _arr(c.vol, 27000, ctypes.c_uint).reshape((30, 30, 30), order="F")
    finally:
        p.close()
    got = (v & DET_MASK) != 0
    want = _expected_detmask(vol.astype(np.uint32), dets)
    assert want.sum() > 20
    assert np.array_equal(got, want)


@pytest.mark.parametrize("flags,media", [("DSPMXVW", 3), ("DPX", 2), ("P", 2), ("DSPMXVW", 5), ("XVW", 4)])
def test_detected_record_length(flags, media):
    """record = detid + (nscat, ppath, mom) x (media-1) + exit pos 3 + dir 3 + w0
    (src/mcx_host.cpp:486-496)"""
    prop = [[0, 0, 1, 1]] + [[0.01, 1, 0, 1.37]] * (media - 1)
    p = pmcxcl.prepare({"nphoton": 10, "vol": np.ones((10, 10, 10), np.uint8), "prop": prop, "srcpos": [5, 5, 0],
                        "srcdir": [0, 0, 1], "tstart": 0, "tend": 1e-9, "tstep": 1e-9, "detpos": [[5, 5, 0, 1]],
                        "issavedet": 1, "savedetflag": flags})
    try:
        m = media - 1
        want = ("D" in flags) + m * (("S" in flags) + ("P" in flags) + ("M" in flags)) + 3 * ("X" in flags) + \
            3 * ("V" in flags) + ("W" in flags)
        assert lib.mcx_detreclen(ctypes.byref(p.c)) == want
    finally:
        p.close()


def test_invalid_config_reports_error():
    base = {"nphoton": 10, "vol": np.ones((10, 10, 10), np.uint8), "prop": [[0, 0, 1, 1], [0.01, 1, 0, 1]],
            "srcpos": [5, 5, 0], "srcdir": [0, 0, 1], "tstart": 0, "tend": 1e-9, "tstep": 1e-9}
    with pytest.raises(ValueError):
        pmcxcl.prepare(dict(base, srctype="no-such-source"))
    # caught by the C validation (src/mcx_utils.c mcx_validatecfg), reported without exiting
    with pytest.raises(RuntimeError):
        pmcxcl.prepare(dict(base, tstart=2e-9))
    with pytest.raises(RuntimeError):
        pmcxcl.prepare(dict(base, vol=np.full((10, 10, 10), 5, np.uint8)))

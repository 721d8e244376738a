"""Packed media formats (src/mcx_core.cl:1079-1193 updateproperty; packing
src/mcx_utils.c:3597-3681 and src/pmcxcl.cpp:243-389).  CPU only.

A packed format stores a voxel's properties in its 32-bit volume word instead of
a label: mua_float (one float), muamus_float/muamus_half (two halves),
labelplus (a label with one component replaced by a half), asgn_byte and
muamus_short (bytes/shorts mixing media 1 and 2).

Pins:
  * the packer against binary16 itself: exactly representable values give
    numpy's float16 bits; other values are truncated (never above the value,
    within one half ulp below); the subnormal path rounds to the nearest half;
    the special words (zero -> EPS / 0x00010000, NaN -> void) are the
    reference's;
  * the oracle's decoding by equivalence: a packed volume whose words carry the
    same properties as a label volume gives the label run's result bit for bit
    (uniform mua_float; halves that are exact; labelplus replacing a component
    by its own or a new exact value; the mixing formats with media 1 = 2);
  * the mixing direction: muamus_short words at full scale reproduce a label
    run of medium 2 to float rounding, at zero scale medium 1;
  * the file reader packs mua_float/muamus_float files as the binding does and
    reads muamus_half/asgn_byte/muamus_short words as stored."""

import ctypes
import json
import os

import numpy as np
import pytest

import oracle
from helpers import sorted_records
from mcxcl_amd import _native, workloads
from test_config import _arr, _readconfig

lib = _native.lib
NT = 4096
MUA_FLOAT, AS_F2H, LABEL_HALF, AS_HALF, ASGN_BYTE, AS_SHORT = 101, 100, 99, 102, 103, 104


def _pack(fmt, vals, unitinmm=1.0):
    vals = np.ascontiguousarray(vals, dtype=np.float32)
    n = vals.size // {MUA_FLOAT: 1, AS_F2H: 2, LABEL_HALF: 3}[fmt]
    out = np.zeros(n, np.uint32)
    rc = lib.mcx_pack_media(fmt, vals.ctypes.data, n, unitinmm, out.ctypes.data)
    return rc, out


def _f32bits(x):
    return int(np.float32(x).view(np.uint32))


def _h16bits(x):
    return int(np.float16(x).view(np.uint16))


# ---------------------------------------------------------------- packing

def test_mua_float_words():
    rc, w = _pack(MUA_FLOAT, [0.01, 0.0, np.nan, -0.0, 0.25, -0.02])
    assert rc == 0
    assert w[0] == _f32bits(0.01)
    assert w[1] == _f32bits(1e-10)          # zero would read as a void voxel: EPS instead
    assert w[2] == 0 and w[3] == 0          # NaN and -0 are void
    assert w[4] == _f32bits(0.25)
    assert w[5] == _f32bits(-0.02)          # the sign bit: the kernel takes |mua|
    rc, w = _pack(MUA_FLOAT, [0.01], unitinmm=0.5)
    assert w[0] == _f32bits(np.float32(0.01) * np.float32(0.5))


def test_half_packing_against_binary16():
    rng = np.random.default_rng(1)
    # exactly representable halves (normal range) come out as numpy's float16 bits
    exact = np.float16(rng.uniform(2.0 ** -14, 60000.0, 400)).astype(np.float32)
    pairs = np.stack([exact[:200], exact[200:]], axis=1)
    rc, w = _pack(AS_F2H, pairs)
    assert rc == 0
    lo, hi = w & 0xFFFF, w >> 16
    assert np.array_equal(lo, np.float16(pairs[:, 0]).view(np.uint16))
    assert np.array_equal(hi, np.float16(pairs[:, 1]).view(np.uint16))
    # other values in the normal range are truncated: h <= x < h + ulp(h)
    x = rng.uniform(2.0 ** -13, 60000.0, 2000).astype(np.float32)
    rc, w = _pack(AS_F2H, np.stack([x, x], axis=1))
    h = (w & 0xFFFF).astype(np.uint16).view(np.float16).astype(np.float64)
    up = np.nextafter(np.float16(h), np.float16(np.inf)).astype(np.float64)
    assert (h <= x).all() and (x < up).all()
    # subnormal halves: values in [2^-24, 2^-15) whose top mantissa bits are small
    # are rounded to the nearest subnormal
    sub = np.float32([2.0 ** -20, 3 * 2.0 ** -24, 2.0 ** -16 * 1.001, 2.0 ** -24 * 5.5])
    rc, w = _pack(AS_F2H, np.stack([sub, sub], axis=1))
    for v, word in zip(sub, w):
        m = (int(np.float32(v).view(np.uint32)) >> 13) & 0x3FF
        if m < 0x10:
            assert abs(float(np.uint16(word & 0xFFFF).view(np.float16)) - float(v)) <= 2.0 ** -25 + 1e-30


def test_half_special_words():
    rc, w = _pack(AS_F2H, [[0.0, 0.0], [np.nan, 1.0], [1.0, np.nan], [0.5, 2.0]])
    assert w[0] == 0x00010000               # an all-zero word would be void: the smallest mus instead
    assert w[1] == 0 and w[2] == 0          # NaN in either component: void
    assert w[3] == (_h16bits(2.0) << 16) | _h16bits(0.5)
    rc, w = _pack(AS_F2H, [[0.5, 2.0]], unitinmm=0.25)
    assert w[0] == (_h16bits(0.5) << 16) | _h16bits(0.125)


def test_labelplus_words():
    rc, w = _pack(LABEL_HALF, [[0.5, 0, 3], [2.0, 1, 7], [0.25, 2, 1], [1.5, 3, 16383]], unitinmm=0.5)
    assert rc == 0
    # {value, component, label}: label | component << 14 low, the value's half high;
    # components 1 and 2 are scaled by the length unit, as the reference's module does
    assert w[0] == (_h16bits(0.5) << 16) | (0 << 14) | 3
    assert w[1] == (_h16bits(1.0) << 16) | (1 << 14) | 7
    assert w[2] == (_h16bits(0.125) << 16) | (2 << 14) | 1
    assert w[3] == (_h16bits(1.5) << 16) | (3 << 14) | 16383
    assert _pack(LABEL_HALF, [[0.5, 4, 3]])[0] == -2
    assert _pack(LABEL_HALF, [[0.5, 1, 16384]])[0] == -2
    assert lib.mcx_pack_media(105, None, 0, 1.0, None) == -1


# ---------------------------------------------------------------- decoding (oracle)

def _run(cfg, **kw):
    return oracle.run(dict(cfg, autopilot=0, nthread=NT, nblocksize=64, rawfield=1, **kw), nthread=NT)


def _same(a, b):
    assert a["stat"]["energyabs"] == b["stat"]["energyabs"]
    assert np.array_equal(a["flux64"], b["flux64"])
    if "detp" in b:
        assert np.array_equal(sorted_records(a["detp"]), sorted_records(b["detp"]))


def _sphere(**kw):
    """60^3 cube (medium 1) with an r=15 sphere of medium 2, exact-half properties;
    detected photons keep their id and exit position (partial paths need labels)"""
    kw.setdefault("savedetflag", "DX")
    kw.setdefault("nphoton", 20000)
    cfg = workloads.sphere_cube_workload(tstep=5e-9, issaveseed=1, **kw)
    cfg["prop"] = [[0, 0, 1, 1], [0.0078125, 1.0, 0.01, 1.37], [0.001953125, 4.0, 0.5, 1.0]]
    return cfg


def test_mua_float_uniform_equals_label_run():
    base = workloads.homogeneous_cube_workload(nphoton=20000, tstep=5e-9)
    prop = np.asarray(base["prop"], np.float32)
    mua = np.full((1,) + base["vol"].shape, prop[1, 0], np.float32)
    _same(_run(dict(base, vol=mua)), _run(base))


@pytest.mark.parametrize("isreflect", [0, 1])
def test_mua_float_nan_voxels_are_void(isreflect):
    """NaN words are void (label-0 rules at their faces), but a packed void voxel
    keeps medium 1's mus and g: the reference writes only mua and n for this
    format (src/mcx_core.cl:1093-1095), so a photon that enters void voxels
    without reflection scatters there as in a label run whose label 0 has
    {0, mus_1, g_1, n_0}"""
    base = workloads.homogeneous_cube_workload(nphoton=20000, tstep=5e-9, isreflect=isreflect)
    base["prop"] = [[0, 1.0, 0.01, 1.0], [0.005, 1.0, 0.01, 1.37]]
    prop = np.asarray(base["prop"], np.float32)
    vol = base["vol"].copy()
    vol[:, :, 40:] = 0                      # a void slab
    mua = np.full((1,) + vol.shape, prop[1, 0], np.float32)
    mua[0][vol == 0] = np.nan
    _same(_run(dict(base, vol=mua)), _run(dict(base, vol=vol)))


def test_muamus_halves_equal_label_run():
    """two media with exact-half mua and mus: g and n of a packed voxel are medium 1's,
    so the sphere gets medium 2's g and n through a label volume only; give both media
    the same g and n"""
    cfg = _sphere()
    cfg["prop"][2][2:] = cfg["prop"][1][2:]
    prop = np.asarray(cfg["prop"], np.float32)
    vol = cfg["vol"]
    mm = np.stack([prop[vol, 0], prop[vol, 1]]).astype(np.float32)
    _same(_run(dict(cfg, vol=mm)), _run(cfg))


def test_labelplus_equals_label_run():
    cfg = _sphere()
    vol = cfg["vol"]
    prop = np.asarray(cfg["prop"], np.float32)
    # every voxel keeps its label and "replaces" mua (component 0) by its own value
    trip = np.zeros((3,) + vol.shape, np.float32)
    trip[2] = vol
    trip[0] = prop[vol, 0]
    trip[1] = 0
    _same(_run(dict(cfg, vol=trip)), _run(cfg))
    # medium-1 voxels replace mus by 2.0 (component 1): a label run with mus_1 = 2.0
    trip[0] = np.where(vol == 1, 2.0, prop[2, 0])
    trip[1] = np.where(vol == 1, 1, 0)
    cfg2 = dict(cfg, prop=[list(r) for r in cfg["prop"]])
    cfg2["prop"][1][1] = 2.0
    _same(_run(dict(cfg, vol=trip)), _run(cfg2))


@pytest.mark.parametrize("fmt", ["asgn_byte", "muamus_short"])
def test_mixing_formats_with_equal_end_members_equal_label_run(fmt):
    """without reflection: asgn_byte mixes n as well, and the background word (0)
    decodes to medium 1 with medium 1's n (src/mcx_core.cl:1162-1166), so with
    reflection on a photon sees no index step at the grid's faces and keeps
    scattering outside until its time gate ends, as in the reference"""
    cfg = workloads.homogeneous_cube_workload(nphoton=20000, tstep=5e-9, isreflect=0)
    cfg["prop"] = [[0, 0, 1, 1], [0.005, 1.0, 0.01, 1.37], [0.005, 1.0, 0.01, 1.37]]
    rng = np.random.default_rng(3)
    shape = cfg["vol"].shape
    if fmt == "asgn_byte":
        vol = rng.integers(1, 127, (4,) + shape).astype(np.int8)
    else:
        vol = rng.integers(1, 32767, (2,) + shape).astype(np.int16)
    _same(_run(dict(cfg, vol=vol)), _run(cfg))


def test_mixing_direction():
    """muamus_short: s/65535 of the way from medium 1 to medium 2"""
    cfg = workloads.homogeneous_cube_workload(nphoton=20000, tstep=5e-9)
    cfg["prop"] = [[0, 0, 1, 1], [0.004, 0.8, 0.01, 1.37], [0.008, 1.6, 0.01, 1.37]]
    shape = cfg["vol"].shape
    full = np.empty((2,) + shape, np.int16)
    full[0] = -1                            # 0xFFFF: mua at full scale
    full[1] = 32767                         # mus at 32767/65535 (the top bit is the detector mask)
    run = _run(dict(cfg, vol=full))
    mus = np.float32(32767) * np.float32(1 / 65535) * np.float32(0.8) + np.float32(0.8)
    ref = _run(dict(cfg, prop=[[0, 0, 1, 1], [0.008, float(mus), 0.01, 1.37]]))
    np.testing.assert_allclose(run["flux64"].sum(), ref["flux64"].sum(), rtol=2e-4)
    np.testing.assert_allclose(run["stat"]["energyabs"], ref["stat"]["energyabs"], rtol=2e-4)


# ---------------------------------------------------------------- file reader

def _json_cfg(tmp_path, fmt, raw, dims):
    (tmp_path / "vol.bin").write_bytes(raw)
    cfg = {"Session": {"ID": "fmt", "Photons": 1000, "Seed": 1},
           "Domain": {"VolumeFile": "vol.bin", "Dim": list(dims), "MediaFormat": fmt, "LengthUnit": 0.5,
                      "OriginType": 1,
                      "Media": [{"mua": 0, "mus": 0, "g": 1, "n": 1}, {"mua": 0.005, "mus": 1, "g": 0.01, "n": 1.37},
                                {"mua": 0.01, "mus": 2, "g": 0.1, "n": 1.37}]},
           "Forward": {"T0": 0, "T1": 5e-9, "Dt": 5e-9},
           "Optode": {"Source": {"Pos": [2, 2, 0], "Dir": [0, 0, 1]}}}
    (tmp_path / "fmt.json").write_text(json.dumps(cfg))
    return _readconfig(str(tmp_path / "fmt.json"))


@pytest.mark.parametrize("fmt,code,ncomp", [("mua_float", MUA_FLOAT, 1), ("muamus_float", AS_F2H, 2)])
def test_file_reader_packs_float_volumes(tmp_path, fmt, code, ncomp):
    dims = (4, 5, 6)
    vals = np.random.default_rng(5).uniform(0.001, 3.0, (ncomp,) + dims).astype(np.float32)
    c = _json_cfg(tmp_path, fmt, np.asfortranarray(vals).ravel(order="F").tobytes(), dims)
    assert c.mediabyte == code
    words = _arr(c.vol, int(np.prod(dims)), ctypes.c_uint)
    rc, want = _pack(code, np.asfortranarray(vals).ravel(order="F"), unitinmm=0.5)
    assert np.array_equal(words, want)
    lib.mcx_clearcfg(ctypes.byref(c))


@pytest.mark.parametrize("fmt,code", [("muamus_half", AS_HALF), ("asgn_byte", ASGN_BYTE), ("muamus_short", AS_SHORT)])
def test_file_reader_keeps_word_formats(tmp_path, fmt, code):
    dims = (4, 5, 6)
    words = np.random.default_rng(6).integers(1, 0x7FFFFFFF, int(np.prod(dims))).astype(np.uint32)
    c = _json_cfg(tmp_path, fmt, words.tobytes(), dims)
    assert c.mediabyte == code
    assert np.array_equal(_arr(c.vol, words.size, ctypes.c_uint), words)
    lib.mcx_clearcfg(ctypes.byref(c))


# ---------------------------------------------------------------- rejected combinations

def test_unsupported_combinations_are_rejected_before_device_work():
    import mcxcl_amd
    cfg = _sphere(nphoton=100)
    vol = cfg["vol"]
    prop = np.asarray(cfg["prop"], np.float32)
    trip = np.stack([prop[vol, 0], np.zeros(vol.shape, np.float32), vol.astype(np.float32)])
    # labelplus words carry a replaced property, not a label to tally partial paths by
    with pytest.raises(RuntimeError, match="label volume"):
        mcxcl_amd.run(dict(cfg, vol=trip, savedetflag="DP"))
    # per-voxel mua/mus arrays go with label volumes only
    mua = prop[vol, 0][None].astype(np.float32)
    with pytest.raises(RuntimeError, match="label volume"):
        mcxcl_amd.run(dict(cfg, vol=mua, muamus=np.stack([prop[vol, 0], prop[vol, 1]]).astype(np.float32)))
    # nor with split-voxel volumes
    sv = np.zeros((8,) + vol.shape, np.uint8)
    sv[0] = vol
    with pytest.raises(RuntimeError, match="label volume"):
        mcxcl_amd.run(dict(cfg, vol=sv, muamus=np.stack([prop[vol, 0], prop[vol, 1]]).astype(np.float32)))
    # mixed-label volumes are not built
    with pytest.raises(ValueError, match="mixlabel"):
        mcxcl_amd.run(dict(cfg, vol=np.zeros((3,) + vol.shape, np.int16)))

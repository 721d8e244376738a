"""Output files and replay seed files (the reference's writers and readers in
src/mcx_utils.c:487-1182 and :3707-3897), through the C ABI.  CPU only.

The byte layouts are checked against the published formats, read here
independently of the writers: the NIfTI-1 header by its field offsets
(nifti1.h), JNIfTI/JData by json + base64 + zlib/gzip, BJData by the Draft 2
markers, and the .mch History header by the 4 + 15 x 4 byte layout that the
reference's own Python reader walks (utils/python/__init__.py:49-80)."""

import base64
import ctypes
import gzip
import json
import os
import struct
import zlib

import numpy as np
import pytest

from mcxcl_amd import _native, files, pmcxcl

lib = _native.lib
DIM = (4, 5, 6)
NG = 3


def _prepared(**extra):
    vol = np.ones(DIM, np.uint8)
    vol[:, :, 3:] = 2
    cfg = dict(nphoton=1000, vol=vol, tstart=0, tend=3e-9, tstep=1e-9, srcpos=[2, 2, 0], srcdir=[0, 0, 1],
               prop=[[0, 0, 1, 1], [0.01, 1.0, 0.9, 1.37], [0.02, 2.0, 0.8, 1.4]], unitinmm=0.5,
               detpos=[[2, 2, 0, 1.5], [1, 1, 0, 1.0]])
    cfg.update(extra)
    return pmcxcl.prepare(cfg)


def _save(c, data, fmt, tmp_path, session="out", zipid=0):
    c.session = session.encode()
    c.rootpath = str(tmp_path).encode()
    c.outputformat = fmt
    c.zipid = zipid
    c.isdumpjson = 1   # no progress lines
    arr = np.ascontiguousarray(data, np.float32)
    rc = lib.mcx_try_savedata(arr.ctypes.data, arr.size, ctypes.byref(c))
    assert rc == 0, _native.last_error()
    return os.path.join(str(tmp_path), session)


def _field():
    rng = np.random.default_rng(7)
    return rng.random(np.prod(DIM) * NG).astype(np.float32)


@pytest.fixture
def prep():
    p = _prepared()
    yield p
    p.close()


def test_raw_formats(prep, tmp_path):
    f = _field()
    base = _save(prep.c, f, 0, tmp_path)                          # mc2
    assert np.array_equal(np.fromfile(base + ".mc2", "<f4"), f)
    base = _save(prep.c, f, 3, tmp_path)                          # ubj: raw floats too
    assert np.array_equal(np.fromfile(base + ".ubj", "<f4"), f)
    base = _save(prep.c, f, 4, tmp_path)                          # tx3
    raw = open(base + ".tx3", "rb").read()
    assert struct.unpack_from("<4I", raw) == (0x8814, *DIM)
    assert np.array_equal(np.frombuffer(raw, "<f4", offset=16), f)
    vol, _ = files.loadfile(base + ".tx3")
    assert np.array_equal(vol, f.reshape(DIM + (NG,), order="F"))


# published nifti1.h byte offsets of the fields the writer sets
NII_OFFSETS = {"sizeof_hdr": 0, "dim": 40, "intent_code": 68, "datatype": 70, "bitpix": 72, "pixdim": 76,
               "vox_offset": 108, "scl_slope": 112, "xyzt_units": 123, "magic": 344}


def test_nifti_reader_layout_is_the_published_one():
    assert files.NII_HEADER.itemsize == 348
    for name, off in NII_OFFSETS.items():
        assert files.NII_HEADER.fields[name][1] == off, name


@pytest.mark.parametrize("fmt,ext", [(1, ".nii"), (2, ".hdr")])
def test_nifti_header_and_data(prep, tmp_path, fmt, ext):
    f = _field()
    base = _save(prep.c, f, fmt, tmp_path)
    raw = open(base + ext, "rb").read()
    assert len(raw) == (348 + 4 + f.nbytes if fmt == 1 else 348)
    assert struct.unpack_from("<i", raw, 0)[0] == 348
    assert struct.unpack_from("<8h", raw, 40) == (4, *DIM, NG, 0, 0, 0)
    assert struct.unpack_from("<3h", raw, 68) == (0, 16, 32)          # intent none, float32, 32 bits
    pix = struct.unpack_from("<8f", raw, 76)
    assert pix[1:4] == (0.5, 0.5, 0.5) and pix[4] == pytest.approx(1e-3)   # tstep in microseconds
    assert struct.unpack_from("<2f", raw, 108) == ((352.0, 0.0) if fmt == 1 else (0.0, 0.0))
    assert raw[123] == 2 | 24                                          # mm, microseconds
    assert raw[344:348] == (b"n+1\0" if fmt == 1 else b"ni1\0")
    data = raw[352:] if fmt == 1 else open(base + ".img", "rb").read()
    assert np.array_equal(np.frombuffer(data, "<f4"), f)
    vol, hdr = files.loadnii(base + ext)
    assert np.array_equal(vol, f.reshape(DIM + (NG,), order="F"))


def test_nifti_label_volume_is_written_as_uint16_pairs(prep, tmp_path):
    labels = np.array([1, 2 | 0x80000000, 0, 7], np.uint32)
    name = os.path.join(str(tmp_path), "mask").encode()
    c = prep.c
    lib.mcx_savenii.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                ctypes.POINTER(_native.Config)]
    lib.mcx_savenii(labels.ctypes.data, labels.size, name, 768, 1, ctypes.byref(c))
    raw = open(name.decode() + ".nii", "rb").read()
    assert struct.unpack_from("<3h", raw, 68)[1:] == (512, 16)
    assert struct.unpack_from("<5h", raw, 40) == (4, 2, *DIM)
    assert np.frombuffer(raw, "<u2", offset=352).tolist() == [1, 0, 2, 1, 0, 0, 7, 0]


@pytest.mark.parametrize("zipid,unzip", [(0, zlib.decompress), (1, gzip.decompress), (2, lambda b: b)])
def test_jnii(prep, tmp_path, zipid, unzip):
    f = _field()
    base = _save(prep.c, f, 5, tmp_path, zipid=zipid)
    doc = json.load(open(base + ".jnii"))
    hdr, dat = doc["NIFTIHeader"], doc["NIFTIData"]
    assert hdr["Dim"] == [*DIM, NG, 1] and hdr["DataType"] == "single" and hdr["NIIHeaderSize"] == 348
    assert hdr["VoxelSize"] == pytest.approx([0.5, 0.5, 0.5, 1e-9, 1])   # voxel edge in mm, tstep in s
    assert hdr["Description"] == "MCX volumetric output: Fluence rate (W/mm^2)"
    assert hdr["Name"] == "out" and hdr["LastSliceID"] == NG
    assert dat["_ArrayType_"] == "single" and dat["_ArraySize_"] == [*DIM, NG, 1]
    assert dat["_ArrayOrder_"] == "c" and dat["_ArrayZipSize_"] == f.size
    assert dat["_ArrayZipType_"] == ["zlib", "gzip", "base64"][zipid]
    assert np.array_equal(np.frombuffer(unzip(base64.b64decode(dat["_ArrayZipData_"])), "<f4"), f)
    vol, _ = files.loadjnii(base + ".jnii")
    assert np.array_equal(vol, f.reshape(DIM + (NG, 1), order="F"))


def _bjdata(buf):
    """A BJData Draft 2 reader written from the specification for this test."""
    num = {b"i": "<i1", b"U": "<u1", b"I": "<i2", b"u": "<u2", b"l": "<i4", b"m": "<u4", b"L": "<i8",
           b"M": "<u8", b"d": "<f4", b"D": "<f8"}
    pos = [0]

    def take(n):
        out = buf[pos[0]:pos[0] + n]
        assert len(out) == n
        pos[0] += n
        return out

    def scalar(m):
        dt = np.dtype(num[m])
        return np.frombuffer(take(dt.itemsize), dt)[0].item()

    def length():
        return scalar(take(1))

    def value(m=None):
        m = m or take(1)
        if m in num:
            return scalar(m)
        if m == b"S":
            return take(length()).decode()
        if m == b"C":
            return take(1).decode()
        if m in (b"[", b"{"):
            typ = cnt = None
            if buf[pos[0]:pos[0] + 1] == b"$":
                take(1)
                typ = take(1)
            if buf[pos[0]:pos[0] + 1] == b"#":
                take(1)
                cnt = length()
            if m == b"[" and typ in num:
                return np.frombuffer(take(np.dtype(num[typ]).itemsize * cnt), num[typ]).tolist()
            out = {} if m == b"{" else []
            end = b"}" if m == b"{" else b"]"
            while (cnt is None and buf[pos[0]:pos[0] + 1] != end) or (cnt is not None and len(out) < cnt):
                if m == b"{":
                    k = take(length()).decode()
                    out[k] = value(typ)
                else:
                    out.append(value(typ))
            if cnt is None:
                take(1)
            return out
        raise AssertionError("unexpected marker %r" % m)

    doc = value()
    assert pos[0] == len(buf)
    return doc


def test_bnii_matches_jnii(prep, tmp_path):
    f = _field()
    base = _save(prep.c, f, 6, tmp_path)
    doc = _bjdata(open(base + ".bnii", "rb").read())
    jbase = _save(prep.c, f, 5, tmp_path, session="ref")
    jdoc = json.load(open(jbase + ".jnii"))
    hdr, jhdr = doc["NIFTIHeader"], jdoc["NIFTIHeader"]
    for key in ("NIIHeaderSize", "Dim", "DataType", "BitDepth", "Unit", "Description", "Affine", "SForm"):
        assert hdr[key] == jhdr[key], key
    assert hdr["VoxelSize"] == pytest.approx(jhdr["VoxelSize"])
    assert hdr["Orientation"] == {"x": "r", "y": "a", "z": "s"}
    dat = doc["NIFTIData"]
    assert dat["_ArraySize_"] == [*DIM, NG, 1] and dat["_ArrayOrder_"] == "c"
    assert np.array_equal(np.frombuffer(zlib.decompress(bytes(dat["_ArrayZipData_"])), "<f4"), f)
    vol, _ = files.loadbnii(base + ".bnii")
    assert np.array_equal(vol, f.reshape(DIM + (NG, 1), order="F"))


def test_unbuilt_codec_is_an_error(prep, tmp_path):
    c = prep.c
    c.session, c.rootpath, c.outputformat, c.zipid, c.isdumpjson = b"x", str(tmp_path).encode(), 5, 4, 1
    f = _field()
    assert lib.mcx_try_savedata(f.ctypes.data, f.size, ctypes.byref(c)) != 0
    assert "compression" in _native.last_error()
    assert not os.path.exists(os.path.join(str(tmp_path), "x.jnii"))


def _records(n, flag, maxmedia=2, seed=3):
    """synthetic detected-photon records in the savedetflag column order"""
    rng = np.random.default_rng(seed)
    cols = []
    if flag & 1:
        cols.append(rng.integers(1, 3, n).astype(np.float32)[:, None])
    if flag & 2:
        cols.append(rng.integers(0, 200, (n, maxmedia)).astype(np.uint32).view(np.float32))
    if flag & 4:
        cols.append(rng.uniform(0, 40, (n, maxmedia)).astype(np.float32))
    if flag & 8:
        cols.append(rng.random((n, maxmedia)).astype(np.float32))
    if flag & 16:
        cols.append(rng.uniform(0, 4, (n, 3)).astype(np.float32))
    if flag & 32:
        cols.append(rng.uniform(-1, 1, (n, 3)).astype(np.float32))
    if flag & 64:
        cols.append(np.ones((n, 1), np.float32))
    return np.ascontiguousarray(np.hstack(cols), np.float32)


def _fill_his(c, rec, nseed):
    c.his.colcount = rec.shape[1]
    c.his.savedphoton = rec.shape[0]
    c.his.detected = rec.shape[0]
    c.his.totalphoton = 1000
    c.his.unitinmm = 0.5
    c.his.seedbyte = 16 if nseed else 0
    c.his.normalizer = 2.5


def _savedet(c, rec, seeds, tmp_path, fmt, doappend=0, session="det"):
    c.session = session.encode()
    c.rootpath = str(tmp_path).encode()
    c.outputformat = fmt
    c.isdumpjson = 1
    c.issaveseed = 1 if seeds is not None else 0
    sp = seeds.ctypes.data if seeds is not None else None
    rc = lib.mcx_try_savedetphoton(rec.ctypes.data, sp, rec.shape[0], doappend, ctypes.byref(c))
    assert rc == 0, _native.last_error()
    return os.path.join(str(tmp_path), session)


def _walk_mch(raw):
    """blocks of (header tuple, records, seeds) by the reference reader's walk"""
    out, off = [], 0
    while off < len(raw):
        assert raw[off:off + 4] == b"MCXH"
        # version maxmedia detnum colcount totalphoton detected savedphoton | unitinmm | seedbyte |
        # normalizer | respin | srcnum savedetflag totalsource | reserved
        h = struct.unpack_from("<7IfIfi3Ii", raw, off + 4)
        off += 64
        maxmedia, colcount, saved, seedbyte = h[1], h[3], h[6], h[8]
        rec = np.frombuffer(raw, "<f4", saved * colcount, off).reshape(saved, colcount)
        off += 4 * saved * colcount
        seeds = np.frombuffer(raw, np.uint8, saved * seedbyte, off).reshape(saved, seedbyte)
        off += saved * seedbyte
        out.append((h, rec, seeds))
    return out


def test_mch_header_records_seeds_and_append(prep, tmp_path):
    c = prep.c
    rec = _records(9, 0x15)
    seeds = np.arange(9 * 16, dtype=np.uint8).reshape(9, 16)
    _fill_his(c, rec, 1)
    base = _savedet(c, rec, seeds, tmp_path, 0)
    rec2 = _records(4, 0x15, seed=5)
    _fill_his(c, rec2, 1)
    _savedet(c, rec2, seeds[:4].copy(), tmp_path, 0, doappend=1)
    blocks = _walk_mch(open(base + ".mch", "rb").read())
    assert len(blocks) == 2
    h, r, s = blocks[0]
    assert h[0] == 1 and h[1] == 2 and h[2] == 2 and h[3] == 1 + 2 + 3   # version, media, detectors, DPX columns
    assert h[4:7] == (1000, 9, 9) and h[7] == 0.5 and h[8] == 16 and h[9] == 2.5
    assert np.array_equal(r, rec) and np.array_equal(s, seeds)
    assert np.array_equal(blocks[1][1], rec2) and np.array_equal(blocks[1][2], seeds[:4])
    data, heads, sd = files.loadmch(base + ".mch")
    assert data.shape == (13, 6) and len(heads) == 2 and sd.shape == (13, 16)


def test_mch_without_seeds_and_trajectory_tag(prep, tmp_path):
    c = prep.c
    rec = _records(5, 0x5)
    _fill_his(c, rec, 0)
    base = _savedet(c, rec, None, tmp_path, 0)
    raw = open(base + ".mch", "rb").read()
    assert len(raw) == 64 + rec.nbytes
    c.his.detected = 0          # trajectory records go to .mct (src/mcx_utils.c:972)
    base = _savedet(c, rec, None, tmp_path, 0, session="traj")
    assert os.path.exists(base + ".mct")


@pytest.mark.parametrize("flag", [0x7F, 0x15, 0x05])
def test_jdat_columns(prep, tmp_path, flag):
    c = prep.c
    c.savedetflag = flag
    c.his.savedetflag = flag
    rec = _records(11, flag)
    seeds = np.random.default_rng(1).integers(0, 256, (11, 16)).astype(np.uint8)
    _fill_his(c, rec, 1)
    base = _savedet(c, rec, seeds, tmp_path, 5)
    doc = json.load(open(base + "_detp.jdat"))["MCXData"]
    info = doc["Info"]
    assert info["ColumnNum"] == rec.shape[1] and info["DetectedPhoton"] == 11 and info["SeedByte"] == 16
    assert info["LengthUnit"] == 0.5 and info["Normalizer"] == 2.5 and info["SaveDetFlag"] == flag
    assert [m["mua"] for m in info["Media"]] == pytest.approx([0, 0.01, 0.02])   # back in 1/mm
    pd = doc["PhotonData"]
    widths = [("detid", 1), ("nscat", 2), ("ppath", 2), ("mom", 2), ("p", 3), ("v", 3), ("w0", 1)]
    col = 0
    for bit, (name, w) in enumerate(widths):
        if not flag >> bit & 1:
            assert name not in pd
            continue
        a = pd[name]
        raw = zlib.decompress(base64.b64decode(a["_ArrayZipData_"]))
        assert a["_ArraySize_"] == [11, w]
        if name == "detid":
            assert a["_ArrayType_"] == "uint32"
            assert np.array_equal(np.frombuffer(raw, "<u4"), rec[:, 0].astype(np.uint32))
        else:
            # nscat is stored as raw uint32 bits in the record and kept as such
            assert np.array_equal(np.frombuffer(raw, "<u4").reshape(11, w), rec[:, col:col + w].view(np.uint32))
        col += w
    assert col == rec.shape[1]
    s = pd["seed"]
    assert s["_ArrayType_"] == "uint8" and s["_ArraySize_"] == [11, 16]
    assert np.array_equal(np.frombuffer(zlib.decompress(base64.b64decode(s["_ArrayZipData_"])), np.uint8),
                          seeds.ravel())


def test_jdat_splits_the_source_index_from_detid(tmp_path):
    p = _prepared()
    try:
        c = p.c
        c.savedetflag = 0x5
        c.his.savedetflag = 0x5
        c.his.totalsource = 3
        rec = _records(6, 0x5)
        src = np.array([1, 2, 3, 1, 2, 3], np.uint32)
        rec[:, 0] = (rec[:, 0].astype(np.uint32) | (src << 16)).astype(np.float32)
        _fill_his(c, rec, 0)
        base = _savedet(c, rec, None, tmp_path, 5)
        out = files.loadjdat(base + "_detp.jdat")
        assert np.array_equal(out["srcid"].ravel(), src)
        assert np.array_equal(out["detid"].ravel(), rec[:, 0].astype(np.uint32) & 0xFFFF)
    finally:
        p.close()


def _parsecmd(args):
    c = _native.Config()
    lib.mcx_initcfg(ctypes.byref(c))
    argv = (ctypes.c_char_p * len(args))(*[a.encode() for a in args])
    rc = lib.mcx_try_parsecmd(len(args), argv, ctypes.byref(c))
    return c, rc


def _json_input(tmp_path, **session):
    cfg = {"Session": dict({"ID": "base", "Photons": 1000, "RNGSeed": 7, "DoSaveSeed": 1}, **session),
           "Forward": {"T0": 0, "T1": 5e-9, "Dt": 5e-9},
           "Domain": {"Media": [{"mua": 0, "mus": 0, "g": 1, "n": 1},
                                {"mua": 0.01, "mus": 1, "g": 0.9, "n": 1.37},
                                {"mua": 0.05, "mus": 2, "g": 0.8, "n": 1.4}],
                      "Dim": [10, 10, 10], "OriginType": 1, "LengthUnit": 0.5},
           "Shapes": [{"Grid": {"Tag": 1, "Size": [10, 10, 10]}}, {"ZLayers": [[6, 10, 2]]}],
           "Optode": {"Source": {"Pos": [5, 5, 0], "Dir": [0, 0, 1]},
                      "Detector": [{"Pos": [3, 5, 0], "R": 2}, {"Pos": [7, 5, 0], "R": 2}]}}
    path = os.path.join(str(tmp_path), "in.json")
    json.dump(cfg, open(path, "w"))
    return path


def _expected_replay(rec, seeds, unit, replaydet, tend=5e-9):
    """weights and times of flight of the replayed photons from the D and P
    columns (reference src/mcx_utils.c:1432-1466, media in 1/mm)"""
    mua = np.array([0.01, 0.05], np.float32)
    n = np.array([1.37, 1.4], np.float32)
    keep, w, tof, det = [], [], [], []
    for i in range(rec.shape[0]):
        d = int(rec[i, 0])
        if replaydet > 0 and d != replaydet:
            continue
        plen = rec[i, 1:3].astype(np.float32) * np.float32(unit)
        wi = np.float32(1)
        ti = np.float32(0)
        for j in range(2):
            wi = np.float32(wi * np.exp(np.float32(-mua[j] * plen[j]), dtype=np.float32))
            ti = np.float32(ti + np.float32(np.float32(plen[j] * np.float32(3.335640951981520e-12)) * n[j]))
        if ti < 0 or ti > tend:
            continue
        keep.append(i)
        w.append(wi)
        tof.append(ti)
        det.append(d)
    return np.array(keep), np.array(w, np.float32), np.array(tof, np.float32), np.array(det)


@pytest.mark.parametrize("fmt,replaydet", [("mch", 0), ("mch", 2), ("jdat", 0), ("jdat", 1)])
def test_replay_seed_files_load_through_the_command_line(tmp_path, fmt, replaydet):
    """-E <file> with -O j: seeds, weights, times of flight and detector ids of
    the saved photons, filtered by -P (reference src/mcx_utils.c:3707-3897)"""
    inp = _json_input(tmp_path)
    base_c, rc = _parsecmd(["mcxcl", "-f", inp])
    assert rc == 0, _native.last_error()
    rec = _records(20, 0x5)
    rec[:, 1:3] *= 0.3
    rec[3, 1:3] = 5000.0            # outside the time window: dropped
    seeds = np.random.default_rng(2).integers(0, 256, (20, 16)).astype(np.uint8)
    base_c.savedetflag = 0x5
    base_c.his.savedetflag = 0x5
    _fill_his(base_c, rec, 1)
    base = _savedet(base_c, rec, seeds, tmp_path, 0 if fmt == "mch" else 5, session="base")
    lib.mcx_clearcfg(ctypes.byref(base_c))
    seedfile = base + (".mch" if fmt == "mch" else "_detp.jdat")
    c, rc = _parsecmd(["mcxcl", "-f", inp, "-E", seedfile, "-O", "j", "-P", str(replaydet)])
    assert rc == 0, _native.last_error()
    try:
        keep, w, tof, det = _expected_replay(rec, seeds, 0.5, replaydet)
        n = len(keep)
        assert c.seed == -999 and c.nphoton == n and 0 < n < 20
        got_seed = np.ctypeslib.as_array((ctypes.c_uint8 * (16 * n)).from_address(c.replay.seed)).reshape(n, 16)
        assert np.array_equal(got_seed, seeds[keep])
        got_w = np.ctypeslib.as_array((ctypes.c_float * n).from_address(c.replay.weight))
        got_t = np.ctypeslib.as_array((ctypes.c_float * n).from_address(c.replay.tof))
        got_d = np.ctypeslib.as_array((ctypes.c_int * n).from_address(c.replay.detid))
        np.testing.assert_allclose(got_w, w, rtol=2e-6)
        np.testing.assert_allclose(got_t, tof, rtol=2e-6)
        assert np.array_equal(got_d & 0xFFFF, det)
    finally:
        lib.mcx_clearcfg(ctypes.byref(c))


def test_plain_replay_keeps_every_saved_photon(tmp_path):
    inp = _json_input(tmp_path)
    base_c, rc = _parsecmd(["mcxcl", "-f", inp])
    assert rc == 0
    rec = _records(7, 0x5)
    rec[2, 1:3] = 5000.0
    seeds = np.arange(7 * 16, dtype=np.uint8).reshape(7, 16)
    base_c.savedetflag = 0x5
    base_c.his.savedetflag = 0x5
    _fill_his(base_c, rec, 1)
    base = _savedet(base_c, rec, seeds, tmp_path, 0, session="plain")
    lib.mcx_clearcfg(ctypes.byref(base_c))
    c, rc = _parsecmd(["mcxcl", "-f", inp, "-E", base + ".mch"])
    assert rc == 0, _native.last_error()
    try:
        assert c.nphoton == 7 and c.seed == -999
        w = np.ctypeslib.as_array((ctypes.c_float * 7).from_address(c.replay.weight))
        assert np.all(w == 1.0)
    finally:
        lib.mcx_clearcfg(ctypes.byref(c))


def test_seed_file_errors(tmp_path):
    inp = _json_input(tmp_path)
    base_c, _ = _parsecmd(["mcxcl", "-f", inp])
    rec = _records(4, 0x5)
    base_c.savedetflag = 0x5
    base_c.his.savedetflag = 0x5
    _fill_his(base_c, rec, 0)
    base = _savedet(base_c, rec, None, tmp_path, 0, session="noseed")
    lib.mcx_clearcfg(ctypes.byref(base_c))
    c, rc = _parsecmd(["mcxcl", "-f", inp, "-E", base + ".mch"])
    assert rc != 0 and "seed data" in _native.last_error()
    lib.mcx_clearcfg(ctypes.byref(c))
    c, rc = _parsecmd(["mcxcl", "-f", inp, "-E", os.path.join(str(tmp_path), "missing.mch")])
    assert rc != 0
    lib.mcx_clearcfg(ctypes.byref(c))


@pytest.mark.parametrize("encoding", ["zlib", "array"])
def test_shapes_given_as_a_jdata_volume(tmp_path, encoding):
    """Shapes holding a JData label array (src/mcx_utils.c:3239-3250): row-major
    [Nx, Ny, Nz] in the file, x fastest once loaded"""
    lab = np.random.default_rng(4).integers(0, 3, (4, 5, 6)).astype(np.uint8)
    if encoding == "zlib":
        shapes = {"_ArrayType_": "uint8", "_ArraySize_": [4, 5, 6], "_ArrayZipType_": "zlib",
                  "_ArrayZipSize_": lab.size,
                  "_ArrayZipData_": base64.b64encode(zlib.compress(lab.tobytes())).decode()}
    else:
        shapes = {"_ArrayType_": "uint8", "_ArraySize_": [4, 5, 6], "_ArrayData_": lab.ravel().tolist()}
    cfg = {"Session": {"ID": "jd", "Photons": 10},
           "Forward": {"T0": 0, "T1": 5e-9, "Dt": 5e-9},
           "Domain": {"Media": [{"mua": 0, "mus": 0, "g": 1, "n": 1}, {"mua": 0.01, "mus": 1, "g": 0.9, "n": 1.37},
                                {"mua": 0.02, "mus": 1, "g": 0.9, "n": 1.37}], "OriginType": 1},
           "Shapes": shapes,
           "Optode": {"Source": {"Pos": [2, 2, 0], "Dir": [0, 0, 1]}}}
    c = _native.Config()
    lib.mcx_initcfg(ctypes.byref(c))
    rc = lib.mcx_try_loadjson_string(json.dumps(cfg).encode(), ctypes.byref(c))
    assert rc == 0, _native.last_error()
    try:
        assert (c.dim.x, c.dim.y, c.dim.z) == (4, 5, 6)
        vol = np.ctypeslib.as_array((ctypes.c_uint * lab.size).from_address(c.vol)).copy()
        assert np.array_equal(vol & 0x7FFFFFFF, lab.ravel(order="F"))
    finally:
        lib.mcx_clearcfg(ctypes.byref(c))


def test_jdata_codec_roundtrip_against_python():
    """the C decoder reads what Python's zlib/gzip/base64 write"""
    for dtype, ztype, z in [("single", "zlib", zlib.compress), ("uint16", "gzip", gzip.compress),
                            ("int32", "base64", lambda b: b)]:
        a = (np.arange(60) * 3 - 50).astype(files._JTYPES[dtype]).reshape(3, 4, 5)
        obj = {"_ArrayType_": dtype, "_ArraySize_": [3, 4, 5], "_ArrayZipType_": ztype,
               "_ArrayZipData_": base64.b64encode(z(a.tobytes())).decode()}
        assert np.array_equal(files.jdata_decode(obj), a)

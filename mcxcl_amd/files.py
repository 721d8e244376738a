"""Readers for the files the simulator writes.

The reference ``pmcxcl`` package re-exports ``loadmc2``, ``loadmch``,
``loadnii`` and ``loadfile`` from the separate ``pmcx`` package
(pmcxcl/pmcxcl/__init__.py:36-66); these are the same readers for the
formats this build writes (mcxcl_amd/csrc/mcx_output.c):

  .mc2 / .ubj     raw float32, x fastest
  .tx3            GL_RGBA32F tag, 3 dims, float32 data
  .nii / .hdr     NIfTI-1 header (+ .img for .hdr)
  .jnii / .bnii   JNIfTI in JSON or BJData with a compressed JData array
  .mch / .mct     64-byte History headers + records (+ seeds), blocks may repeat
  _detp.jdat      JSON MCXData with one JData array per record column group

Only numpy, json, zlib and base64 are used.
"""

import base64
import gzip
import json
import struct
import zlib

import numpy as np

__all__ = ["loadmc2", "loadmch", "loadnii", "loadjnii", "loadbnii", "loadjdat", "loadfile",
           "jdata_decode", "bjdata_loads"]

_JTYPES = {"int8": np.int8, "uint8": np.uint8, "int16": np.int16, "uint16": np.uint16,
           "int32": np.int32, "uint32": np.uint32, "int64": np.int64, "uint64": np.uint64,
           "single": np.float32, "float32": np.float32, "double": np.float64, "float64": np.float64}

# History header of .mch/.mct files (reference src/mcx_utils.h:113-130)
_HIS = np.dtype([("magic", "S4"), ("version", "<u4"), ("maxmedia", "<u4"), ("detnum", "<u4"),
                 ("colcount", "<u4"), ("totalphoton", "<u4"), ("detected", "<u4"),
                 ("savedphoton", "<u4"), ("unitinmm", "<f4"), ("seedbyte", "<u4"),
                 ("normalizer", "<f4"), ("respin", "<i4"), ("srcnum", "<u4"),
                 ("savedetflag", "<u4"), ("totalsource", "<u4"), ("reserved", "<i4")])


def loadmc2(path, dims):
    """float32 volume of shape ``dims`` (x fastest), e.g. [Nx, Ny, Nz, Ngate]."""
    return np.fromfile(path, dtype="<f4").reshape(dims, order="F")


def loadmch(path):
    """(records [N, colcount] float32, list of header dicts, seeds [N, seedbyte] uint8 or None).

    Blocks written with append mode are concatenated; path lengths stay in
    grid units (multiply by header['unitinmm'] for mm)."""
    raw = open(path, "rb").read()
    off, recs, seeds, heads = 0, [], [], []
    while off < len(raw):
        h = np.frombuffer(raw, _HIS, 1, off)[0]
        if h["magic"] != b"MCXH":
            raise ValueError("%s: block at byte %d is not an MCXH header" % (path, off))
        off += _HIS.itemsize
        n, c, sb = int(h["savedphoton"]), int(h["colcount"]), int(h["seedbyte"])
        recs.append(np.frombuffer(raw, "<f4", n * c, off).reshape(n, c))
        off += 4 * n * c
        if sb and off + n * sb <= len(raw):
            # seeds follow the records when the run saved them (issaveseed)
            nxt = off + n * sb
            if nxt == len(raw) or raw[nxt:nxt + 4] == b"MCXH":
                seeds.append(np.frombuffer(raw, np.uint8, n * sb, off).reshape(n, sb))
                off = nxt
        heads.append({k: h[k].item() for k in _HIS.names if k not in ("magic", "reserved")})
    data = np.concatenate(recs) if recs else np.zeros((0, 0), np.float32)
    return data, heads, (np.concatenate(seeds) if seeds else None)


def _nii_header():
    # NIfTI-1 header fields in order (nifti1.h); offsets follow from the sizes
    return np.dtype([("sizeof_hdr", "<i4"), ("data_type", "S10"), ("db_name", "S18"), ("extents", "<i4"),
                     ("session_error", "<i2"), ("regular", "S1"), ("dim_info", "u1"), ("dim", "<i2", 8),
                     ("intent_p1", "<f4"), ("intent_p2", "<f4"), ("intent_p3", "<f4"), ("intent_code", "<i2"),
                     ("datatype", "<i2"), ("bitpix", "<i2"), ("slice_start", "<i2"), ("pixdim", "<f4", 8),
                     ("vox_offset", "<f4"), ("scl_slope", "<f4"), ("scl_inter", "<f4"), ("slice_end", "<i2"),
                     ("slice_code", "u1"), ("xyzt_units", "u1"), ("cal_max", "<f4"), ("cal_min", "<f4"),
                     ("slice_duration", "<f4"), ("toffset", "<f4"), ("glmax", "<i4"), ("glmin", "<i4"),
                     ("descrip", "S80"), ("aux_file", "S24"), ("qform_code", "<i2"), ("sform_code", "<i2"),
                     ("quatern", "<f4", 3), ("qoffset", "<f4", 3), ("srow", "<f4", 12),
                     ("intent_name", "S16"), ("magic", "S4")])


NII_HEADER = _nii_header()
_NII_TYPES = {2: np.uint8, 4: np.int16, 8: np.int32, 16: np.float32, 64: np.float64, 256: np.int8,
              512: np.uint16, 768: np.uint32}


def loadnii(path):
    """(volume, header record) of a .nii or .hdr/.img pair; volume x fastest."""
    raw = open(path, "rb").read()
    h = np.frombuffer(raw, NII_HEADER, 1, 0)[0]
    if h["sizeof_hdr"] != 348:
        raise ValueError("%s: not a NIfTI-1 header" % path)
    nd = int(h["dim"][0])
    shape = [int(d) for d in h["dim"][1:1 + nd]]
    dt = np.dtype(_NII_TYPES[int(h["datatype"])]).newbyteorder("<")
    if path.endswith(".hdr"):
        data = open(path[:-4] + ".img", "rb").read()
        off = 0
    else:
        data, off = raw, int(h["vox_offset"])
    vol = np.frombuffer(data, dt, int(np.prod(shape)), off)
    return vol.reshape(shape, order="F"), h


def _unzip(ztype, blob):
    if ztype == "zlib":
        return zlib.decompress(blob)
    if ztype == "gzip":
        return gzip.decompress(blob)
    if ztype == "base64":
        return blob
    raise ValueError("JData codec %r is not supported" % ztype)


def jdata_decode(obj):
    """numpy array of a JData annotated array (_ArrayZipData_ or _ArrayData_)."""
    dt = np.dtype(_JTYPES[obj["_ArrayType_"]]).newbyteorder("<")
    size = obj["_ArraySize_"]
    size = [int(size)] if np.isscalar(size) else [int(s) for s in size]
    if "_ArrayZipData_" in obj:
        blob = obj["_ArrayZipData_"]
        blob = base64.b64decode(blob) if isinstance(blob, str) else bytes(np.asarray(blob, np.uint8))
        flat = np.frombuffer(_unzip(obj.get("_ArrayZipType_", "zlib"), blob), dt)
    else:
        flat = np.asarray(obj["_ArrayData_"], dt).ravel()
    order = "F" if obj.get("_ArrayOrder_", "r") in ("c", "col", "column") else "C"
    return flat.reshape(size, order=order)


def loadjnii(path):
    """(volume, NIFTIHeader dict) of a .jnii file."""
    doc = json.load(open(path))
    return jdata_decode(doc["NIFTIData"]), doc["NIFTIHeader"]


# ---- BJData Draft 2 (little-endian UBJSON derivative) ----
_BJ_NUM = {b"i": "<i1", b"U": "<u1", b"I": "<i2", b"u": "<u2", b"l": "<i4", b"m": "<u4", b"L": "<i8",
           b"M": "<u8", b"h": "<f2", b"d": "<f4", b"D": "<f8"}


class _BJReader:
    def __init__(self, buf):
        self.b, self.o = buf, 0

    def take(self, n):
        out = self.b[self.o:self.o + n]
        if len(out) != n:
            raise ValueError("truncated BJData")
        self.o += n
        return out

    def marker(self):
        return self.take(1)

    def number(self, m):
        dt = np.dtype(_BJ_NUM[m])
        return np.frombuffer(self.take(dt.itemsize), dt)[0].item()

    def length(self):
        return int(self.number(self.marker()))

    def value(self, m=None):
        m = m or self.marker()
        if m in _BJ_NUM:
            return self.number(m)
        if m == b"S" or m == b"H":
            return self.take(self.length()).decode()
        if m == b"C":
            return self.take(1).decode()
        if m == b"T":
            return True
        if m == b"F":
            return False
        if m == b"Z":
            return None
        if m == b"[":
            return self.container(b"]", False)
        if m == b"{":
            return self.container(b"}", True)
        raise ValueError("unknown BJData marker %r at %d" % (m, self.o - 1))

    def container(self, end, isobj):
        typ, count = None, None
        if self.b[self.o:self.o + 1] == b"$":
            self.o += 1
            typ = self.marker()
        if self.b[self.o:self.o + 1] == b"#":
            self.o += 1
            count = self.length()
        if typ is not None and count is not None and not isobj and typ in _BJ_NUM:
            dt = np.dtype(_BJ_NUM[typ])
            return np.frombuffer(self.take(dt.itemsize * count), dt).copy()
        out = {} if isobj else []
        i = 0
        while True:
            if count is None:
                if self.b[self.o:self.o + 1] == end:
                    self.o += 1
                    return out
            elif i == count:
                return out
            if isobj:
                key = self.take(self.length()).decode()
                out[key] = self.value(typ)
            else:
                out.append(self.value(typ))
            i += 1


def bjdata_loads(buf):
    """Decode one BJData value (bytes) into dicts, lists, numpy vectors and scalars."""
    return _BJReader(bytes(buf)).value()


def loadbnii(path):
    """(volume, NIFTIHeader dict) of a .bnii file."""
    doc = bjdata_loads(open(path, "rb").read())
    return jdata_decode(doc["NIFTIData"]), doc["NIFTIHeader"]


def loadjdat(path):
    """dict of the MCXData.Info fields and the decoded PhotonData columns."""
    doc = json.load(open(path))["MCXData"]
    out = {"info": doc["Info"]}
    for k, v in doc.get("PhotonData", {}).items():
        out[k] = jdata_decode(v)
    return out


def loadfile(path, dims=None):
    """Dispatch on the extension (reference pmcx.loadfile): volumes return
    (data, header), detected-photon files their reader's result."""
    low = path.lower()
    if low.endswith(".mc2") or low.endswith(".ubj"):
        if dims is None:
            raise ValueError("raw .mc2 data needs dims")
        return loadmc2(path, dims), None
    if low.endswith(".tx3"):
        raw = open(path, "rb").read()
        tag, nx, ny, nz = struct.unpack_from("<4I", raw, 0)
        data = np.frombuffer(raw, "<f4", offset=16)
        return data.reshape([nx, ny, nz, -1], order="F"), {"tag": tag}
    if low.endswith(".nii") or low.endswith(".hdr"):
        return loadnii(path)
    if low.endswith(".jnii"):
        return loadjnii(path)
    if low.endswith(".bnii"):
        return loadbnii(path)
    if low.endswith(".mch") or low.endswith(".mct"):
        return loadmch(path)
    if low.endswith(".jdat"):
        return loadjdat(path)
    raise ValueError("unknown file type: %s" % path)

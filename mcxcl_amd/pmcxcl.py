"""pmcxcl-compatible Python interface: ``run(cfg)``, ``gpuinfo()``, ``version()``.

Mirrors the reference binding (src/pmcxcl.cpp): the same configuration keys
(parse_config, src/pmcxcl.cpp:428-1117), the same volume dtype rules
(parseVolume, :93-427), the same call sequence (initcfg -> validatecfg ->
run_simulation -> clearcfg, :1124-1553) and the same outputs: ``flux``
(Fortran-ordered ``[Nx, Ny, Nz, Nt(, Ndet)]`` float32), ``detp``
(``[reclen, ndetected]``), ``seeds`` (``[16, ndetected]`` uint8), ``dref``,
``vol`` (with ``dumpmask``) and ``stat``.  Errors raise ``ValueError`` /
``RuntimeError`` as the reference binding does.
"""

import ctypes

import numpy as np

from . import _native
from ._native import lib, Config, GPUInfo

SRCTYPES = ["pencil", "isotropic", "cone", "gaussian", "planar", "pattern", "fourier", "arcsine", "disk",
            "fourierx", "fourierx2d", "zgaussian", "line", "slit", "pencilarray", "pattern3d", "hyperboloid", "ring"]
# pmcxcl output-type names (src/pmcxcl.cpp:854-877); indices >= 5 shift down by 2
OUTPUTTYPES = ["flux", "fluence", "energy", "jacobian", "nscat", "wl", "wp", "wm", "rf", "length", "rfmus",
               "wltof", "wptof", "adjoint", "adjoint_dcoeff", "adjoint_mus", "adjoint_musp", "adjoint_mua_d",
               "adjoint_mua_musp"]
SAVEFLAGS = "DSPMXVWI"
# first adjoint output type (enum TOutputType, src/mcx_utils.h:58-62) and the
# output keys of each adjoint type (src/pmcxcl.cpp:1426-1456)
OT_ADJOINT = 11
OT_RF, OT_RFMUS = 6, 8
ADJOINT_KEYS = {11: ("jmua",), 12: ("jd",), 13: ("jmus",), 14: ("jmusp",), 15: ("jmua", "jd"), 16: ("jmua", "jmusp")}
DEBUGFLAGS = "RMP"
SEED_FROM_FILE = -999

_INT_FIELDS = ["nphoton", "nblocksize", "nthread", "maxdetphoton", "maxgate", "respin", "isreflect", "isref3",
               "isrefint", "isnormalized", "isgpuinfo", "issrcfrom0", "autopilot", "printnum", "voidtime",
               "issaveref", "issaveexit", "ismomentum", "isspecular", "istrajstokes", "replaydet", "faststep",
               "maxvoidstep", "maxjumpdebug", "internalsrc", "gscatter", "srcnum", "srcid", "optlevel", "isatomic",
               "issave2pt", "issavedet", "issaveseed", "mathmode"]
_FLOAT_FIELDS = ["tstart", "tstep", "tend", "sradius", "minenergy", "unitinmm", "omega", "lambda"]


def _set_vec(dst, values, n):
    vals = [float(v) for v in np.asarray(values, dtype=np.float64).ravel()[:n]]
    for i, v in enumerate(vals):
        setattr(dst, "xyzw"[i], v)


def _carray(arr, dtype):
    """Copy a numpy array into C-allocated memory (freed by mcx_clearcfg)."""
    a = np.ascontiguousarray(arr, dtype=dtype)
    ptr = _native.cmalloc(a.nbytes)
    ctypes.memmove(ptr, a.ctypes.data, a.nbytes)
    return ptr


# packed media formats (include/mcx_utils.h)
MEDIA_LABEL_HALF, MEDIA_AS_F2H, MEDIA_MUA_FLOAT, MEDIA_ASGN_BYTE, MEDIA_AS_SHORT = 99, 100, 101, 103, 104
MEDIA_2LABEL_SPLIT = 97
MCX_DEBUG_MOVE = 2
MCX_DEBUG_MOVE_ONLY = 8


def _pack_float_media(vol, fmt, unitinmm):
    """float32 [k, Nx, Ny, Nz] per-voxel values to one packed word per voxel (the
    library's mcx_pack_media, the packing of src/pmcxcl.cpp:243-389)"""
    src = np.ascontiguousarray(np.asfortranarray(vol).ravel(order="F"), dtype=np.float32)
    out = np.empty(int(np.prod(vol.shape[1:])), dtype=np.uint32)     # voxels, x fastest
    rc = _native.lib.mcx_pack_media(fmt, src.ctypes.data, out.size, unitinmm, out.ctypes.data)
    if rc:
        raise ValueError("the 2nd volume must have an integer value between 0 and 3, and labels below 16384"
                         if rc == -2 else "unsupported packed media format %d" % fmt)
    return out


def _parse_volume(user, c):
    """reference parseVolume (src/pmcxcl.cpp:93-427): dtype and leading dimension
    select the media format.  3-D integer arrays are label volumes; 4-D arrays are
    per-voxel properties packed into one word per voxel: float32 [1,...] mua
    (mua_float), float32 [2,...] {mua, mus} as two halves (muamus_float),
    float32 [3,...] {value, component, label} (labelplus), int8 [4,...] four
    mixing bytes (asgn_byte), int16 [2,...] two mixing shorts (muamus_short).
    uint8/int8 [8,...] are split voxels (svmc): {lower label, upper label,
    reference point x, y, z, normal x, y, z} per voxel, passed as read (two
    words per voxel) and laid out by the library.  Per-voxel float32 {mua, mus}
    without half rounding go in 'muamus' beside a label volume."""
    if "vol" not in user:
        raise ValueError("Configuration must specify a 2/3/4D volume.")
    vol = np.asarray(user["vol"])
    if vol.ndim == 4:
        k = vol.shape[0]
        if vol.dtype == np.float32 and k in (1, 2, 3):
            c.mediabyte = {1: MEDIA_MUA_FLOAT, 2: MEDIA_AS_F2H, 3: MEDIA_LABEL_HALF}[k]
            words = _pack_float_media(vol, c.mediabyte, float(c.unitinmm))
        elif vol.dtype == np.int8 and k == 4:
            c.mediabyte = MEDIA_ASGN_BYTE
            words = np.ascontiguousarray(np.asfortranarray(vol).ravel(order="F")).view(np.uint32)
        elif vol.dtype in (np.int8, np.uint8) and k == 8:
            c.mediabyte = MEDIA_2LABEL_SPLIT
            words = np.ascontiguousarray(np.asfortranarray(vol).ravel(order="F")).view(np.uint32)
        elif vol.dtype == np.int16 and k == 2:
            c.mediabyte = MEDIA_AS_SHORT
            words = np.ascontiguousarray(np.asfortranarray(vol).ravel(order="F")).view(np.uint32)
        else:
            raise ValueError("4-D volumes must be float32 [1|2|3, Nx, Ny, Nz], int8 [4, Nx, Ny, Nz], uint8/int8 "
                             "[8, Nx, Ny, Nz] or int16 [2, Nx, Ny, Nz] (the mixlabel format is not supported)")
        c.dim.x, c.dim.y, c.dim.z = (int(d) for d in vol.shape[1:])
        c.vol = _carray(words.astype(np.uint32), np.uint32)
        return
    if vol.ndim != 3:
        raise ValueError("the 'vol' field must be a 3-D array")
    if vol.dtype in (np.uint8, np.int8):
        c.mediabyte = 1
    elif vol.dtype in (np.uint16, np.int16):
        c.mediabyte = 2
    elif vol.dtype in (np.uint32, np.int32):
        c.mediabyte = 4
    else:
        raise ValueError("unsupported volume data type %s (use an integer label volume)" % vol.dtype)
    c.dim.x, c.dim.y, c.dim.z = (int(d) for d in vol.shape)
    flat = np.asfortranarray(vol).ravel(order="F")
    flat = flat.view(np.uint8).astype(np.uint32) if flat.dtype == np.int8 else flat
    flat = flat.view(np.uint16).astype(np.uint32) if flat.dtype == np.int16 else flat
    c.vol = _carray(flat.astype(np.uint32), np.uint32)


def _parse_srcvec(user, c, key):
    """srcpos/srcdir/srcparam1/srcparam2: one row, or one row per source (extra sources)"""
    a = np.atleast_2d(np.asarray(user[key], dtype=np.float32))
    if a.shape[1] < 3 and key in ("srcpos", "srcdir"):
        raise ValueError("the '%s' field must have 3 or 4 columns" % key)
    if a.shape[1] > 4:
        raise ValueError("the '%s' field must have at most 4 columns" % key)
    _set_vec(getattr(c, key), a[0], a.shape[1])
    if a.shape[0] > 1 or c.extrasrclen > 0:
        nextra = a.shape[0] - 1
        if c.extrasrclen and c.extrasrclen != nextra:
            raise ValueError("Length of sub-elements of srcpos/srcdir/srcparam1/srcparam2 must match")
        c.extrasrclen = nextra
        if not c.srcdata:
            c.srcdata = _native.cmalloc(64 * max(1, nextra))
        base = (ctypes.c_float * (16 * nextra)).from_address(c.srcdata)
        slot = {"srcpos": 0, "srcdir": 4, "srcparam1": 8, "srcparam2": 12}[key]
        for i in range(nextra):
            for j in range(a.shape[1]):
                base[i * 16 + slot + j] = float(a[i + 1, j])


def _parse_config(user, c):
    """reference parse_config (src/pmcxcl.cpp:428-1117)"""
    for k in _INT_FIELDS:
        if k in user:
            v = int(user[k])
            if isinstance(getattr(c, k), bytes):
                setattr(c, k, bytes([v & 0xFF]))
            else:
                setattr(c, k, v)
    for k in _FLOAT_FIELDS:
        if k in user:
            setattr(c, k, float(user[k]))
    for k in ("steps", "crop0", "crop1"):
        if k in user:
            v = np.asarray(user[k]).ravel()
            for i, name in enumerate("xyz"):
                setattr(getattr(c, k), name, type(getattr(getattr(c, k), name))(v[i]))
    if "srciquv" in user:
        _set_vec(c.srciquv, user["srciquv"], 4)
    _parse_volume(user, c)
    for key in ("srcpos", "srcdir", "srcparam1", "srcparam2"):
        if key in user:
            _parse_srcvec(user, c, key)
    if "detpos" in user:
        d = np.atleast_2d(np.asarray(user["detpos"], dtype=np.float32))
        if d.size and d.shape[1] != 4:
            raise ValueError("the 'detpos' field must have 4 columns (x,y,z,radius)")
        c.detnum = d.shape[0] if d.size else 0
        if c.detnum:
            c.detpos = _carray(d, np.float32)
    if "detdir" in user:
        d = np.atleast_2d(np.asarray(user["detdir"], dtype=np.float32))
        if d.shape[1] != 4:
            raise ValueError("the 'detdir' field must have 4 columns (vx,vy,vz,focallength)")
        if c.detnum and d.shape[0] != c.detnum:
            raise ValueError("the number of rows in 'detdir' must match 'detpos'")
        c.detdir = _carray(d, np.float32)
    if "prop" in user:
        p = np.atleast_2d(np.asarray(user["prop"], dtype=np.float32))
        if p.shape[1] != 4:
            raise ValueError("the 'prop' field must have 4 columns (mua,mus,g,n)")
        c.medianum = p.shape[0]
        c.prop = _carray(p, np.float32)
    if "polprop" in user:
        # polarised media (src/pmcxcl.cpp:754-785): one row per medium 1.., columns
        # {mua, sphere radius (um), density (1/um^3), n_sphere, n_background}
        pp = np.atleast_2d(np.asarray(user["polprop"], dtype=np.float32))
        if pp.ndim != 2:
            raise ValueError("the 'polprop' field must a 2D array")
        if pp.shape[1] != 5:
            raise ValueError("the 'polprop' field must have 5 columns (mua, radius, rho, n_sph,n_bkg)")
        c.polmedianum = pp.shape[0]
        c.polprop = _carray(pp, np.float32)
    if "session" in user:
        s = str(user["session"]).encode()
        if len(s) >= 64:
            raise ValueError("the 'session' field is too long")
        c.session = s
    if "srctype" in user:
        st = str(user["srctype"]).lower()
        if st not in SRCTYPES:
            raise ValueError("the specified source type is not supported")
        c.srctype = bytes([SRCTYPES.index(st)])
    if "outputtype" in user:
        ot = str(user["outputtype"]).lower()
        if ot not in OUTPUTTYPES:
            raise ValueError("the specified output type is not supported")
        idx = OUTPUTTYPES.index(ot)
        if idx >= 5:
            idx -= 2
        c.outputtype = bytes([idx])
    if "debuglevel" in user:
        lvl = 0
        for ch in str(user["debuglevel"]).upper():
            if ch in DEBUGFLAGS:
                lvl |= 1 << DEBUGFLAGS.index(ch)
        if lvl == 0:
            raise ValueError("the specified debuglevel is not supported")
        c.debuglevel = lvl
    if "savedetflag" in user:
        v = user["savedetflag"]
        if isinstance(v, str):
            flag = 0
            for ch in v.upper():
                if ch in SAVEFLAGS:
                    flag |= 1 << SAVEFLAGS.index(ch)
            c.savedetflag = flag
        else:
            c.savedetflag = int(v)
    if "invcdf" in user:
        v = np.asarray(user["invcdf"], dtype=np.float32).ravel()
        if np.any(np.diff(v) < 0) or np.any(v > 1) or np.any(v < -1):
            raise ValueError("cfg.invcdf contains invalid data; it must be a monotonically increasing vector "
                             "with all values between -1 and 1")
        full = np.concatenate([[-1.0], v, [1.0]]).astype(np.float32)
        c.nphase = full.size
        c.invcdf = _carray(full, np.float32)
    if "angleinvcdf" in user:
        # launch polar angle / pi at equal CDF steps (src/pmcxcl.cpp:963-983)
        v = np.asarray(user["angleinvcdf"], dtype=np.float32).ravel(order="F")
        if np.any(np.diff(v) < 0) or np.any(v > 1) or np.any(v < 0):
            raise ValueError("cfg.angleinvcdf contains invalid data; it must be a monotonically increasing vector "
                             "with all values between 0 and 1")
        c.nangle = v.size
        c.angleinvcdf = _carray(v, np.float32)
    if "srcpattern" in user:
        # pattern weights, Fortran order (pattern index fastest when srcnum > 1; src/pmcxcl.cpp:917-936)
        c.srcpattern = _carray(np.asarray(user["srcpattern"], dtype=np.float32).ravel(order="F"), np.float32)
    if "shapes" in user:
        s = str(user["shapes"]).encode()
        ptr = _native.cmalloc(len(s) + 2)
        ctypes.memmove(ptr, s, len(s))
        c.shapedata = ptr
    if "bc" in user:
        s = str(user["bc"]).encode()
        if not s or len(s) > 12:
            raise ValueError("the 'bc' field must be a non-empty string of at most 12 characters")
        c.bc = s
    if "seed" in user:
        sv = user["seed"]
        if np.isscalar(sv):
            c.seed = int(sv)
        else:
            a = np.asarray(sv, dtype=np.uint8)
            if a.ndim != 2 or a.shape[0] != 16:
                raise ValueError("the row number of cfg.seed does not match RNG seed byte-length")
            c.replay.seed = _carray(np.asfortranarray(a).ravel(order="F"), np.uint8)
            c.seed = SEED_FROM_FILE
            c.nphoton = a.shape[1]
    if "gpuid" in user:
        gid = int(user["gpuid"])
        if gid <= 0 or gid >= 256:
            raise ValueError("GPU id must be positive and can not be more than 256")
        c.gpuid = gid
    if "workload" in user:
        w = np.asarray(user["workload"], dtype=np.float32).ravel()
        for i, v in enumerate(w[:256]):
            c.workload[i] = float(v)


class PreparedConfig:
    """A validated Config (reference parse_config + mcx_validatecfg) that owns its
    C memory; release it with ``close()`` (mcx_clearcfg)."""

    def __init__(self, user):
        self.user = user
        self.c = Config()
        lib.mcx_initcfg(ctypes.byref(self.c))
        self.c.parentid = 2  # mpPython
        self._keep = []
        try:
            _parse_config(user, self.c)
            self._set_pervoxel(user)
            detps_ptr, dims, seedbyte = None, None, 0
            if "detphotons" in user:
                dp = np.asfortranarray(np.asarray(user["detphotons"], dtype=np.float32))
                self._keep.append(dp)
                detps_ptr = dp.ctypes.data
                dims = (ctypes.c_int * 2)(dp.shape[0], dp.shape[1])
                seedbyte = 16 if self.c.replay.seed else 0
            rc = lib.mcx_try_validatecfg(ctypes.byref(self.c), detps_ptr, dims, seedbyte)
            if rc:
                # the reference binding turns errors of the C layer into RuntimeError
                # (src/pmcxcl.cpp:1523-1525)
                raise RuntimeError(_native.last_error())
            c0 = self.c
            if ((c0.outputtype[0] >= OT_ADJOINT or c0.srcid == -2) and c0.seed != SEED_FROM_FILE and c0.detnum > 0
                    and c0.detdir):
                # adjoint outputs and srcid -2: detectors become disk sources (src/pmcxcl.cpp:1179-1198)
                lib.mcx_append_detector_sources(ctypes.byref(c0))
            if not self.c.vol or self.c.medianum == 0:
                raise ValueError("You must define 'vol' and 'prop' field.")
        except Exception:
            self.close()
            raise
        c = self.c
        self.ngate = int((c.tend - c.tstart) / c.tstep + 0.5)
        self.nvox = c.dim.x * c.dim.y * c.dim.z
        self.ndet = c.detnum if (c.seed == SEED_FROM_FILE and c.replaydet == -1) else 1
        self.nsrc = (c.extrasrclen + 1) if (c.extrasrclen and c.srcid < 0) else 1
        # several patterns share each photon: one interleaved slot per pattern
        self.npattern = c.srcnum if (c.srctype[0] in (5, 15) and c.srcnum > 1) else 1
        # complex outputs (RF replay, RF forward): the imaginary half after the real one
        self.nrf = 2 if _is_complex(c) else 1
        self.fieldlen = self.nvox * self.ngate * self.ndet * self.nsrc * self.npattern * self.nrf
        self.raw = None
        self.raw32 = None
        if user.get("rawfield"):
            # unnormalised double-precision accumulator, returned as "flux64"
            self.raw = np.zeros(self.fieldlen, dtype=np.float64)
            c.exportfield64 = self.raw.ctypes.data
        if user.get("rawfield32"):
            # the same accumulator converted to float32 on the device, returned as "flux32"
            self.raw32 = np.empty(self.fieldlen, dtype=np.float32)
            c.exportfield32 = self.raw32.ctypes.data
        self._set_detweight(user)

    def _set_pervoxel(self, user):
        """per-voxel float32 media ('muamus' beside a label volume), replay weights ('replayweight') and exit-weight output ('savedetweight')"""
        c = self.c
        mm = None
        if "muamus" in user:
            mm = np.asarray(user["muamus"], dtype=np.float32)
        if mm is not None:
            if mm.shape != (2, c.dim.x, c.dim.y, c.dim.z):
                raise ValueError("per-voxel mua/mus must be float32 [2, Nx, Ny, Nz] matching the volume")
            if (mm < 0).any() or not np.isfinite(mm).all():
                raise ValueError("per-voxel mua/mus must be finite and non-negative")
            flat = np.asfortranarray(mm).ravel(order="F")   # mua, mus interleaved per voxel
            self._keep.append(flat)
            c.muamus = flat.ctypes.data
        if "replayweight" in user:
            w = np.ascontiguousarray(np.asarray(user["replayweight"], dtype=np.float32).ravel())
            self._keep.append(w)
            c.replayinweight = w.ctypes.data

    def _set_detweight(self, user):
        self.detw = None
        if user.get("savedetweight") and self.c.issavedet:
            self.detw = np.zeros(max(1, int(self.c.maxdetphoton)), dtype=np.float32)
            self.c.exportdetweight = self.detw.ctypes.data

    def collect(self):
        out = _collect(self.c, self.user, self.ngate, self.ndet, self.nsrc)
        if self.detw is not None:
            out["detw"] = self.detw[:self.c.detectedcount].copy()
        dims = [self.c.dim.x, self.c.dim.y, self.c.dim.z, -1]
        if self.raw is not None:
            out["flux64"] = self.raw.reshape(dims, order="F")   # the array the run filled, no copy
        if self.raw32 is not None:
            out["flux32"] = self.raw32.reshape(dims, order="F")
        return out

    def close(self):
        if self.c is not None:
            self.c.exportfield64 = None
            self.c.exportfield32 = None
            self.c.muamus = None
            self.c.exportdetweight = None
            self.c.replayinweight = None
            lib.mcx_clearcfg(ctypes.byref(self.c))
            self.c = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def prepare(cfg=None, **kwargs):
    user = dict(cfg or {})
    user.update(kwargs)
    return PreparedConfig(user)


def run(cfg=None, **kwargs):
    """Run one simulation; ``cfg`` is a dict of pmcxcl keys (or pass them as kwargs).

    Keys beyond the reference's: ``mathmode`` (0: hardware transcendentals,
    1: bit-reproducible arithmetic), ``rawfield`` (also return the unnormalised
    double-precision accumulator as ``flux64``), ``muamus`` (float32
    [2, Nx, Ny, Nz] per-voxel mua/mus in 1/mm beside a label ``vol``),
    ``savedetweight`` (return each detected photon's exit weight as ``detw``),
    ``replayweight`` (replay weights per row of ``detphotons``, required with
    per-voxel media), ``rawfield32`` (the same accumulator converted to float32
    on the device, returned as ``flux32`` at half the read-back bytes) and
    ``rawonly`` (with ``rawfield`` or ``rawfield32``: return only those,
    skipping the normalised float32 copy of large replay grids)."""
    with prepare(cfg, **kwargs) as pc:
        c = pc.c
        if c.issave2pt and not ((pc.user.get("rawfield") or pc.user.get("rawfield32")) and pc.user.get("rawonly")):
            c.exportfield = _native.cmalloc(4 * pc.fieldlen)
        rc = lib.mcx_try_run_simulation(ctypes.byref(c))
        if rc:
            raise RuntimeError("MCX terminated due to an exception: " + _native.last_error())
        return pc.collect()


class Batch:
    """Many simulations in one persistent launch (include/mcx_host.h mcx_batch_*).

    ``Batch(cfgs)`` validates and stages every pmcxcl dict in ``cfgs`` (fast
    arithmetic, no detected photons, no replay); ``run(nstep)`` runs nstep
    steps, each simulating every instance afresh; ``collect()`` returns one
    output dict per instance (``flux``, ``stat``, and ``flux64`` with
    ``rawfield``) from the last step.  ``blocksize`` fixes the lanes per
    workgroup (0: the shape that keeps the most lanes resident); ``probe``
    balances the per-XCD instance lists by costs measured in a probe step
    (else by photon counts).  ``info`` holds the chosen schedule and the last
    step's device time."""

    def __init__(self, cfgs, blocksize=0, probe=True, usegraph=True, nqueue=0):
        self.pcs = []
        self.handle = ctypes.c_void_p()
        try:
            for cfg in cfgs:
                pc = prepare(cfg)
                self.pcs.append(pc)
                if pc.c.issave2pt:
                    pc.c.exportfield = _native.cmalloc(4 * pc.fieldlen)
            n = len(self.pcs)
            self._arr = (ctypes.POINTER(Config) * n)(*[ctypes.pointer(pc.c) for pc in self.pcs])
            self.opt = _native.MCXBatchOpt(int(blocksize), int(not probe), int(bool(usegraph)), int(nqueue))
            rc = lib.mcx_batch_create(self._arr, n, ctypes.byref(self.opt), ctypes.byref(self.handle))
            if rc:
                raise RuntimeError("MCX terminated due to an exception: " + _native.last_error())
        except Exception:
            self.close()
            raise

    @property
    def info(self):
        o = self.opt
        return {"blocksize": o.blocksize, "probe": not o.noprobe, "nblock": o.nblock, "nqueue": o.nqueue,
                "stepms": o.stepms, "kernelms": o.kernelms, "tunems": o.tunems, "usegraph": o.usegraph}

    def run(self, nstep=1):
        rc = lib.mcx_batch_run(self.handle, int(nstep))
        if rc:
            raise RuntimeError("MCX terminated due to an exception: " + _native.last_error())
        return self.info

    def collect(self):
        rc = lib.mcx_batch_collect(self.handle)
        if rc:
            raise RuntimeError("MCX terminated due to an exception: " + _native.last_error())
        return [pc.collect() for pc in self.pcs]

    def close(self):
        if self.handle:
            lib.mcx_batch_destroy(self.handle)
            self.handle = ctypes.c_void_p()
        for pc in self.pcs:
            pc.close()
        self.pcs = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def run_batch(cfgs, nstep=1, **opts):
    """Run the simulations of ``cfgs`` in one persistent launch; returns
    (list of output dicts, schedule info).  See ``Batch``."""
    with Batch(cfgs, **opts) as b:
        b.run(nstep)
        return b.collect(), b.info


def _is_complex(c):
    """complex outputs: RF replay (outputtype r/s) or RF forward (omega > 0 without replay)"""
    return (c.seed == SEED_FROM_FILE and c.outputtype[0] in (OT_RF, OT_RFMUS)) or \
        (c.omega > 0 and c.seed != SEED_FROM_FILE)


def _collect(c, user, ngate, ndet, nsrc):
    out = {}
    reclen = lib.mcx_detreclen(ctypes.byref(c))
    n = c.detectedcount
    if c.issaveseed and c.seeddata and c.issavedet:
        seeds = np.ctypeslib.as_array((ctypes.c_uint8 * (16 * n)).from_address(c.seeddata)).copy()
        out["seeds"] = seeds.reshape(n, 16).T.copy(order="F")
    if user.get("dumpmask"):
        vol = np.ctypeslib.as_array((ctypes.c_uint32 * (c.dim.x * c.dim.y * c.dim.z)).from_address(c.vol)).copy()
        out["vol"] = vol.reshape((c.dim.x, c.dim.y, c.dim.z), order="F")
    if c.issavedet and n > 0 and reclen > 0:
        det = np.ctypeslib.as_array((ctypes.c_float * (reclen * n)).from_address(c.exportdetected)).copy()
        out["detp"] = np.asfortranarray(det.reshape(n, reclen).T)
    if (c.debuglevel & (MCX_DEBUG_MOVE | MCX_DEBUG_MOVE_ONLY)) and c.exportdebugdata and c.debugdatalen:
        # trajectory positions [6, N]: photon id (uint32 bits), x, y, z, weight, source id
        # (src/pmcxcl.cpp:1266-1283)
        ts = c.istrajstokes
        nrow = 6 + 4 * int(ts[0] if isinstance(ts, bytes) else ts)     # Stokes vectors follow with istrajstokes
        traj = np.ctypeslib.as_array((ctypes.c_float * (nrow * c.debugdatalen)).from_address(c.exportdebugdata)).copy()
        out["traj"] = np.asfortranarray(traj.reshape(c.debugdatalen, nrow).T)
    if c.issave2pt and c.exportfield:
        dims = [c.srcnum * c.dim.x, c.dim.y, c.dim.z, ngate]
        nrf = 2 if _is_complex(c) else 1
        if ndet > 1:
            dims.append(ndet)
        if nsrc > 1 or nrf > 1:
            # RF replay: real and imaginary halves along the last axis (src/pmcxcl.cpp:1339-1345)
            dims += [1] * (5 - len(dims)) + [nsrc * nrf]
        total = int(np.prod(dims))
        flux = np.ctypeslib.as_array((ctypes.c_float * total).from_address(c.exportfield)).copy()
        flux = flux.reshape(dims, order="F")
        if c.issaveref:
            volmask = np.ctypeslib.as_array(
                (ctypes.c_uint32 * (c.dim.x * c.dim.y * c.dim.z)).from_address(c.vol)).copy()
            inside = (volmask & 0x7FFFFFFF).reshape((c.dim.x, c.dim.y, c.dim.z), order="F") > 0
            fl = flux.reshape((c.dim.x, c.dim.y, c.dim.z, -1), order="F")
            dref = np.where(inside[..., None], 0.0, -fl).astype(np.float32)
            fl[~inside] = 0.0
            out["dref"] = dref.reshape(dims, order="F")
            flux = fl.reshape(dims, order="F")
        if c.omega > 0 and c.seed != SEED_FROM_FILE:
            # RF forward: complex fluence, real half + 1j * imaginary half (src/pmcxcl.cpp:1387-1408);
            # with pooled sources one complex field per source along the last axis
            n = flux.shape[-1] // 2
            flux = (flux[..., :n] + 1j * flux[..., n:]).astype(np.complex64)
            if n == 1:
                flux = flux[..., 0]
                if flux.ndim == 5 and flux.shape[4] == 1:
                    flux = flux[..., 0]
        out["flux"] = flux
    ot = c.outputtype[0]
    if ot >= OT_ADJOINT and c.exportjacob and c.detnum and c.extrasrclen + 1 > c.detnum:
        # adjoint Jacobians [Nx, Ny, Nz, Ns*Nd], pair s*Nd+d (src/pmcxcl.cpp:1414-1491)
        npair = (c.extrasrclen + 1 - c.detnum) * c.detnum
        adjlen = c.dim.x * c.dim.y * c.dim.z * npair
        keys = ADJOINT_KEYS.get(ot, ())
        # RF forward: complex64, the imaginary blocks after the real ones (src/pmcxcl.cpp:1459-1485)
        nrf = 2 if (c.omega > 0 and c.seed != SEED_FROM_FILE) else 1
        buf = np.ctypeslib.as_array((ctypes.c_float * (adjlen * len(keys) * nrf)).from_address(c.exportjacob)).copy()
        for k, name in enumerate(keys):
            v = buf[k * adjlen:(k + 1) * adjlen]
            if nrf == 2:
                im = buf[(len(keys) + k) * adjlen:(len(keys) + k + 1) * adjlen]
                v = (v + 1j * im).astype(np.complex64)
            out[name] = v.reshape((c.dim.x, c.dim.y, c.dim.z, npair), order="F")
    out["stat"] = {
        "runtime": int(c.runtime),
        "nphoton": int(c.nphoton) * (c.respin if c.respin > 1 else 1),
        "energytot": float(c.energytot),
        "energyabs": float(c.energyabs),
        "normalizer": float(c.normalizer),
        "unitinmm": float(c.unitinmm),
        "workload": [1.0],
        "kernelms": float(c.kernelms),
        "jobms": float(c.jobms),
        "runtimems": float(c.runtimems),
        "detected": int(c.his.detected),
    }
    if c.polprop and c.polmedianum:
        # the optical properties the polarised run used, per mm (src/pmcxcl.cpp:1511-1521)
        prop = np.ctypeslib.as_array((ctypes.c_float * (4 * c.medianum)).from_address(c.prop)).copy()
        prop = prop.reshape(c.medianum, 4)
        prop[1:c.polmedianum + 1, :2] /= np.float32(c.unitinmm)
        out["prop"] = np.asfortranarray(prop.T)
    return out


def gpuinfo():
    """List the HIP devices (reference pmcxcl.gpuinfo, src/pmcxcl.cpp:1591-1646)."""
    c = Config()
    lib.mcx_initcfg(ctypes.byref(c))
    active = ctypes.c_uint(0)
    devs = (ctypes.c_int * 256)()
    info = ctypes.POINTER(GPUInfo)()
    n = lib.mcx_list_gpu(ctypes.byref(c), ctypes.byref(active), devs, ctypes.byref(info))
    res = []
    for i in range(n):
        g = info[i]
        res.append({k: (getattr(g, k).decode() if k == "name" else getattr(g, k))
                    for k, _ in GPUInfo._fields_})
    if n:
        lib.mcx_cleargpuinfo(ctypes.byref(info))
    lib.mcx_clearcfg(ctypes.byref(c))
    return res


def version():
    return lib.mcx_version_string().decode()

"""Loader for the native library ``libmcxhip.so`` and a ctypes mirror of ``Config``.

The library is built in-tree (``make -C mcxcl_amd``, or ``__graft_entry__.build()``)
into ``mcxcl_amd/lib/``.  There is no fallback: importing this module without the
built library raises ``ImportError``.

The ctypes ``Config`` structure is laid out from the offsets the library reports
(``mcx_config_offsetof``), so the mirror cannot drift from ``include/mcx_utils.h``.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "lib", "libmcxhip.so")

if not os.path.exists(LIBPATH):
    raise ImportError(
        "mcxcl_amd native library not found at %s; build it with `make -C mcxcl_amd` "
        "or __graft_entry__.build()" % LIBPATH
    )

lib = ctypes.CDLL(LIBPATH, mode=ctypes.RTLD_GLOBAL)
libc = ctypes.CDLL("libc.so.6")
libc.malloc.restype = ctypes.c_void_p
libc.malloc.argtypes = [ctypes.c_size_t]
libc.calloc.restype = ctypes.c_void_p
libc.calloc.argtypes = [ctypes.c_size_t, ctypes.c_size_t]
libc.free.restype = None
libc.free.argtypes = [ctypes.c_void_p]


class Float4(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("z", ctypes.c_float), ("w", ctypes.c_float)]


class UInt4(ctypes.Structure):
    _fields_ = [("x", ctypes.c_uint), ("y", ctypes.c_uint), ("z", ctypes.c_uint), ("w", ctypes.c_uint)]


class Medium(ctypes.Structure):
    _fields_ = [("mua", ctypes.c_float), ("mus", ctypes.c_float), ("g", ctypes.c_float), ("n", ctypes.c_float)]


class History(ctypes.Structure):
    _fields_ = [
        ("magic", ctypes.c_char * 4),
        ("version", ctypes.c_uint),
        ("maxmedia", ctypes.c_uint),
        ("detnum", ctypes.c_uint),
        ("colcount", ctypes.c_uint),
        ("totalphoton", ctypes.c_uint),
        ("detected", ctypes.c_uint),
        ("savedphoton", ctypes.c_uint),
        ("unitinmm", ctypes.c_float),
        ("seedbyte", ctypes.c_uint),
        ("normalizer", ctypes.c_float),
        ("respin", ctypes.c_int),
        ("srcnum", ctypes.c_uint),
        ("savedetflag", ctypes.c_uint),
        ("totalsource", ctypes.c_uint),
        ("reserved", ctypes.c_int * 1),
    ]


class Replay(ctypes.Structure):
    _fields_ = [
        ("detid", ctypes.c_void_p),
        ("seed", ctypes.c_void_p),
        ("weight", ctypes.c_void_p),
        ("tof", ctypes.c_void_p),
    ]


class GPUInfo(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 64),
        ("id", ctypes.c_int),
        ("devcount", ctypes.c_int),
        ("platformid", ctypes.c_int),
        ("major", ctypes.c_int),
        ("minor", ctypes.c_int),
        ("globalmem", ctypes.c_size_t),
        ("constmem", ctypes.c_size_t),
        ("sharedmem", ctypes.c_size_t),
        ("regcount", ctypes.c_int),
        ("clock", ctypes.c_int),
        ("sm", ctypes.c_int),
        ("core", ctypes.c_int),
        ("autoblock", ctypes.c_size_t),
        ("autothread", ctypes.c_size_t),
        ("maxgate", ctypes.c_uint),
        ("maxmpthread", ctypes.c_int),
        ("iscpu", ctypes.c_int),
        ("vendor", ctypes.c_int),
    ]


_C = ctypes.c_char
# (name, ctype) in declaration order of struct MCXConfig
CONFIG_FIELDS = [
    ("nphoton", ctypes.c_size_t), ("nblocksize", ctypes.c_uint), ("nthread", ctypes.c_uint), ("seed", ctypes.c_int),
    ("srcpos", Float4), ("srcdir", Float4), ("srciquv", Float4),
    ("tstart", ctypes.c_float), ("tstep", ctypes.c_float), ("tend", ctypes.c_float), ("steps", Float4),
    ("dim", UInt4), ("crop0", UInt4), ("crop1", UInt4),
    ("medianum", ctypes.c_uint), ("polmedianum", ctypes.c_uint), ("detnum", ctypes.c_uint),
    ("maxdetphoton", ctypes.c_uint), ("detradius", ctypes.c_float), ("sradius", ctypes.c_float),
    ("prop", ctypes.c_void_p), ("polprop", ctypes.c_void_p), ("detpos", ctypes.c_void_p),
    ("detdir", ctypes.c_void_p), ("smatrix", ctypes.c_void_p),
    ("maxgate", ctypes.c_uint), ("respin", ctypes.c_int), ("printnum", ctypes.c_uint), ("gpuid", ctypes.c_int),
    ("vol", ctypes.c_void_p), ("session", _C * 64),
    ("isrowmajor", _C), ("isreflect", _C), ("isref3", _C), ("isrefint", _C), ("isnormalized", _C),
    ("issavedet", _C), ("issave2pt", _C), ("isgpuinfo", _C), ("isspecular", _C), ("isverbose", _C),
    ("issrcfrom0", _C), ("ismomentum", _C), ("internalsrc", _C), ("isdumpmask", _C), ("issaveseed", _C),
    ("istrajstokes", _C), ("issaveexit", _C), ("isatomic", _C), ("issaveref", _C), ("isdumpjson", _C),
    ("zipid", ctypes.c_int), ("srctype", _C), ("autopilot", _C), ("outputtype", _C), ("outputformat", _C),
    ("minenergy", ctypes.c_float), ("unitinmm", ctypes.c_float), ("omega", ctypes.c_float),
    ("lambda", ctypes.c_float), ("flog", ctypes.c_void_p), ("his", History),
    ("energytot", ctypes.c_double), ("energyabs", ctypes.c_double), ("energyesc", ctypes.c_double),
    ("rootpath", _C * 1024), ("kernelfile", _C * 64), ("compileropt", _C * 1024),
    ("shapedata", ctypes.c_void_p), ("extrajson", ctypes.c_void_p), ("clsource", ctypes.c_void_p),
    ("maxvoidstep", ctypes.c_int), ("voidtime", ctypes.c_int), ("srcparam1", Float4), ("srcparam2", Float4),
    ("srcnum", ctypes.c_uint), ("srcpattern", ctypes.c_void_p), ("deviceid", _C * 256),
    ("workload", ctypes.c_float * 256), ("exportfield", ctypes.c_void_p), ("exportjacob", ctypes.c_void_p),
    ("exportdetected", ctypes.c_void_p), ("debuglevel", ctypes.c_uint), ("faststep", _C),
    ("normalizer", ctypes.c_float), ("replay", Replay), ("seeddata", ctypes.c_void_p), ("replaydet", ctypes.c_int),
    ("seedfile", _C * 1024), ("jsonfile", _C * 1024), ("maxjumpdebug", ctypes.c_uint),
    ("debugdatalen", ctypes.c_uint), ("gscatter", ctypes.c_uint), ("exportdebugdata", ctypes.c_void_p),
    ("savedetflag", ctypes.c_uint), ("detectedcount", ctypes.c_uint), ("runtime", ctypes.c_uint),
    ("parentid", ctypes.c_int), ("optlevel", ctypes.c_uint), ("mediabyte", ctypes.c_uint), ("bc", _C * 13),
    ("nphase", ctypes.c_uint), ("invcdf", ctypes.c_void_p), ("nangle", ctypes.c_uint),
    ("angleinvcdf", ctypes.c_void_p), ("srcid", ctypes.c_int), ("extrasrclen", ctypes.c_uint),
    ("srcdata", ctypes.c_void_p), ("mathmode", ctypes.c_int), ("kernelms", ctypes.c_float),
    ("jobms", ctypes.c_float), ("runtimems", ctypes.c_double), ("exportfield64", ctypes.c_void_p),
    ("muamus", ctypes.c_void_p), ("exportdetweight", ctypes.c_void_p), ("replayinweight", ctypes.c_void_p),
    ("exportfield32", ctypes.c_void_p),
]

lib.mcx_config_sizeof.restype = ctypes.c_size_t
lib.mcx_config_offsetof.restype = ctypes.c_long
lib.mcx_config_offsetof.argtypes = [ctypes.c_char_p]


def _build_config_type():
    fields = []
    pos = 0
    npad = 0
    for name, ctype in CONFIG_FIELDS:
        off = lib.mcx_config_offsetof(name.encode())
        if off < 0:
            raise ImportError("libmcxhip does not know Config field %r" % name)
        if off < pos:
            raise ImportError("Config mirror out of order at field %r" % name)
        if off > pos:
            fields.append(("_pad%d" % npad, ctypes.c_ubyte * (off - pos)))
            npad += 1
        fields.append((name, ctype))
        pos = off + ctypes.sizeof(ctype)
    total = lib.mcx_config_sizeof()
    if total < pos:
        raise ImportError("Config mirror larger than the C struct")
    if total > pos:
        fields.append(("_pad%d" % npad, ctypes.c_ubyte * (total - pos)))
    return type("Config", (ctypes.Structure,), {"_pack_": 1, "_fields_": fields})


Config = _build_config_type()
assert ctypes.sizeof(Config) == lib.mcx_config_sizeof()

_P = ctypes.POINTER
lib.mcx_initcfg.argtypes = [_P(Config)]
lib.mcx_append_detector_sources.argtypes = [_P(Config)]
lib.mcx_initcfg.restype = None
lib.mcx_clearcfg.argtypes = [_P(Config)]
lib.mcx_clearcfg.restype = None
lib.mcx_try_validatecfg.argtypes = [_P(Config), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
lib.mcx_try_validatecfg.restype = ctypes.c_int
lib.mcx_try_loadjson_string.argtypes = [ctypes.c_char_p, _P(Config)]
lib.mcx_try_loadjson_string.restype = ctypes.c_int
lib.mcx_try_loadbenchmark.argtypes = [ctypes.c_char_p, _P(Config)]
lib.mcx_try_loadbenchmark.restype = ctypes.c_int
lib.mcx_try_readconfig.argtypes = [ctypes.c_char_p, _P(Config)]
lib.mcx_try_readconfig.restype = ctypes.c_int
lib.mcx_try_parse_shapestring.argtypes = [_P(Config), ctypes.c_char_p]
lib.mcx_try_parse_shapestring.restype = ctypes.c_int
lib.mcx_try_parsecmd.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), _P(Config)]
lib.mcx_try_parsecmd.restype = ctypes.c_int
lib.mcx_try_savedata.argtypes = [ctypes.c_void_p, ctypes.c_size_t, _P(Config)]
lib.mcx_try_savedata.restype = ctypes.c_int
lib.mcx_try_savedetphoton.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _P(Config)]
lib.mcx_try_savedetphoton.restype = ctypes.c_int
lib.mcx_pack_media.argtypes = [ctypes.c_uint, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_float, ctypes.c_void_p]
lib.mcx_pack_media.restype = ctypes.c_int
lib.mcx_mie.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
lib.mcx_mie.restype = ctypes.c_int
lib.mcx_try_run_simulation.argtypes = [_P(Config)]
lib.mcx_try_run_simulation.restype = ctypes.c_int
lib.mcx_last_error.restype = ctypes.c_char_p
lib.mcx_list_gpu.argtypes = [_P(Config), _P(ctypes.c_uint), _P(ctypes.c_int), _P(_P(GPUInfo))]
lib.mcx_list_gpu.restype = ctypes.c_int
lib.mcx_device_count.restype = ctypes.c_int
lib.mcx_host_seed_words.argtypes = [ctypes.c_uint, ctypes.c_size_t, ctypes.c_void_p]
lib.mcx_host_seed_words.restype = None
lib.mcx_device_fpmath.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_size_t]
lib.mcx_device_fpmath.restype = ctypes.c_int
lib.mcx_device_rng_stream.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_uint, ctypes.c_void_p]
lib.mcx_device_rng_stream.restype = ctypes.c_int
lib.mcx_device_atomic_ceiling.argtypes = [ctypes.c_int, ctypes.c_uint, ctypes.POINTER(ctypes.c_double)]
lib.mcx_device_atomic_ceiling.restype = ctypes.c_int
lib.mcx_detreclen.argtypes = [_P(Config)]
lib.mcx_detreclen.restype = ctypes.c_uint
lib.mcx_partialdata.argtypes = [_P(Config)]
lib.mcx_partialdata.restype = ctypes.c_uint
lib.mcx_version_string.restype = ctypes.c_char_p
lib.mcx_cleargpuinfo.argtypes = [_P(_P(GPUInfo))]
lib.mcx_cleargpuinfo.restype = None


class MCXBatchOpt(ctypes.Structure):
    """include/mcx_host.h MCXBatchOpt"""
    _fields_ = [("blocksize", ctypes.c_uint), ("noprobe", ctypes.c_uint), ("usegraph", ctypes.c_uint),
                ("nqueue", ctypes.c_uint), ("nblock", ctypes.c_uint), ("reserved", ctypes.c_uint),
                ("stepms", ctypes.c_float), ("kernelms", ctypes.c_float), ("tunems", ctypes.c_float)]


lib.mcx_batch_create.argtypes = [ctypes.POINTER(_P(Config)), ctypes.c_uint, _P(MCXBatchOpt),
                                 ctypes.POINTER(ctypes.c_void_p)]
lib.mcx_batch_create.restype = ctypes.c_int
lib.mcx_batch_run.argtypes = [ctypes.c_void_p, ctypes.c_uint]
lib.mcx_batch_run.restype = ctypes.c_int
lib.mcx_batch_collect.argtypes = [ctypes.c_void_p]
lib.mcx_batch_collect.restype = ctypes.c_int
lib.mcx_batch_destroy.argtypes = [ctypes.c_void_p]
lib.mcx_batch_destroy.restype = None
lib.mcx_run_batch.argtypes = [ctypes.POINTER(_P(Config)), ctypes.c_uint, _P(MCXBatchOpt)]
lib.mcx_run_batch.restype = ctypes.c_int


def last_error():
    msg = lib.mcx_last_error()
    return msg.decode(errors="replace") if msg else ""


def cmalloc(nbytes, zero=True):
    """Allocate with the C allocator so mcx_clearcfg can free it."""
    ptr = libc.calloc(max(1, nbytes), 1) if zero else libc.malloc(max(1, nbytes))
    if not ptr:
        raise MemoryError("malloc(%d) failed" % nbytes)
    return ptr


def cfree(ptr):
    if ptr:
        libc.free(ptr)

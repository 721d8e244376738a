"""The C-ABI library: it loads, exports every function the public headers
declare, and its Config layout agrees with the Python mirror.  No device work."""

import ctypes
import os
import re
import subprocess

import numpy as np

from mcxcl_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("mcx_host.h", "mcx_utils.h")]


def _declared_functions():
    names = []
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \t\*]*?\b(mcx_\w+)\s*\([^;{]*\)\s*;", text, flags=re.M):
            names.append(m.group(1))
    return names


def _exported():
    out = subprocess.run(["nm", "-D", "--defined-only", _native.LIBPATH], stdout=subprocess.PIPE, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_headers_declare_the_drop_in_entry_points():
    names = set(_declared_functions())
    for must in ("mcx_run_simulation", "mcx_list_gpu", "mcx_initcfg", "mcx_clearcfg", "mcx_readconfig",
                 "mcx_loadjson_string", "mcx_validatecfg", "mcx_savedata", "mcx_savedetphoton", "mcx_normalize",
                 "mcx_cleargpuinfo", "mcx_flush", "mcx_error", "mcx_parsecmd"):
        assert must in names, must


def test_library_exports_every_declared_function():
    exported = _exported()
    missing = [n for n in _declared_functions() if n not in exported]
    assert not missing, missing
    for n in _declared_functions():
        assert hasattr(_native.lib, n)


def test_config_layout_matches_mirror():
    assert ctypes.sizeof(_native.Config) == _native.lib.mcx_config_sizeof()
    for name, _ in _native.Config._fields_:
        if name.startswith("_pad"):
            continue
        off = _native.lib.mcx_config_offsetof(name.encode())
        assert off == getattr(_native.Config, name).offset, name


def test_errors_are_returned_not_fatal():
    cfg = _native.Config()
    _native.lib.mcx_initcfg(ctypes.byref(cfg))
    rc = _native.lib.mcx_try_loadjson_string(b'{"Session": {"Photons": 10}, "Domain": ', ctypes.byref(cfg))
    assert rc != 0
    assert _native.last_error()
    _native.lib.mcx_clearcfg(ctypes.byref(cfg))


def test_seed_words_are_glibc_rand():
    """srand(1); rand() x 3 of glibc (TYPE_3 additive generator)"""
    w = np.zeros(3, dtype=np.uint32)
    _native.lib.mcx_host_seed_words(1, 3, w.ctypes.data)
    assert w.tolist() == [1804289383, 846930886, 1681692777]


def test_device_count_without_device():
    """the query works on a host without a GPU (it reports what it sees)"""
    assert _native.lib.mcx_device_count() >= 0

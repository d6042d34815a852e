"""The C-ABI library loads and exports every function include/*.h declares.
No compute calls: this runs without a GPU."""
import ctypes
import glob
import os
import re

from tests.conftest import ROOT


def _declared():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(tarang_\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_library_exports_all_declared_symbols():
    from tarang_amd import _lib
    L = _lib.load()
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from tarang_amd import _lib
    names = set(_declared())
    bound = set(_lib.SIGNATURES) | {"tarang_last_error"}
    assert names == bound, (names - bound, bound - names)


def test_error_channel_without_device():
    from tarang_amd import _lib
    L = _lib.load()
    major, minor = ctypes.c_int(), ctypes.c_int()
    assert L.tarang_version(ctypes.byref(major), ctypes.byref(minor)) == 0
    # NULL arguments are refused with a message, not a crash
    assert L.tarang_version(None, None) == _lib.ERR_INVALID
    assert b"NULL" in L.tarang_last_error()
    assert L.tarang_fivp_step(None, 1) == _lib.ERR_INVALID
    assert L.tarang_forward_transform(None, None, None) == _lib.ERR_INVALID

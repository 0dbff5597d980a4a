"""CPU-side checks of the C-ABI library: it loads, exports every entry point
declared in include/*.h, and rejects bad arguments before touching a GPU."""
import ctypes
import os
import re

import pytest

import cupss_amd as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(inc, f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(cupss_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_entry_point():
    L = C.lib()
    names = declared_functions()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert L.cupss_abi_version() == 1


@pytest.mark.parametrize("dims", [(3, 1, 1), (1, 1, 1), (8192, 1, 1), (16, 12, 1), (16, 16, 5)])
def test_unsupported_grids_are_rejected(dims):
    h = ctypes.c_void_p()
    rc = C.lib().cupss_system_create(dims[0], dims[1], dims[2], 1.0, 1.0, 1.0, 0.1, ctypes.byref(h))
    assert rc == -2
    assert b"power" in C.lib().cupss_last_error() or b"grid" in C.lib().cupss_last_error()


def test_null_arguments_are_rejected():
    L = C.lib()
    assert L.cupss_system_create(16, 16, 1, 1.0, 1.0, 1.0, 0.1, None) == -1
    assert L.cupss_system_prepare(None) == -1
    assert L.cupss_system_advance(None, 1) == -1
    assert L.cupss_system_destroy(None) == 0

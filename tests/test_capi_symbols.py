"""The C-ABI library loads without a GPU and exports every entry point that
include/*.h declares (no compute calls here)."""
import ctypes
import glob
import os
import re

from mapp_amd.dmd import LIB_PATH

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        names |= set(re.findall(r"^\s*(?:int|const char\*|void)\s+\*?(mapp_dmd_\w+)\s*\(", txt, re.M))
    return names


def test_library_exports_every_declared_symbol():
    names = _declared()
    assert len(names) >= 25
    lib = ctypes.CDLL(LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_create_without_gpu_reports_error_not_crash():
    from mapp_amd.dmd import load_library
    lib = load_library()
    ctx = ctypes.c_void_p()
    rc = lib.mapp_dmd_create(ctypes.byref(ctx), 2, 1, 0)
    if rc == 0:  # a GPU is present: nothing to check here
        lib.mapp_dmd_destroy(ctx)
        return
    assert rc < 0 and not ctx.value
    assert lib.mapp_dmd_create(ctypes.byref(ctx), 5, 1, 0) == -4  # E_RANGE: DMD takes nt <= 4
    assert lib.mapp_dmd_create(ctypes.byref(ctx), 3, 3, 0) == -4  # E_RANGE: plain EAM takes nt <= 2

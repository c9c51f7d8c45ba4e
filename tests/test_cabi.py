"""The C ABI library loads and exports every symbol include/*.h declares (CPU)."""

import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))
LIB = os.path.join(ROOT, "dedalus-1.0_amd", "dedalus", "libddx.so")


def declared():
    text = "".join(open(h).read() for h in HEADERS)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ddx_[a-z0-9_]+)\s*\(", text)))


def test_header_parses():
    names = declared()
    assert "ddx_init" in names and "ddx_step" in names and "ddx_conv3x3" in names
    assert len(HEADERS) >= 2 and len(names) >= 30


@pytest.mark.skipif(not os.path.exists(LIB), reason="libddx.so not built")
def test_library_exports_declared_symbols():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    from dedalus import _hip
    assert sorted(_hip.SIGNATURES) == declared()


@pytest.mark.skipif(not os.path.exists(LIB), reason="libddx.so not built")
def test_errors_without_init():
    """Calls that need a device fail with DDX_ERR_NODEV before ddx_init, no crash."""
    from dedalus import _hip
    lib = _hip.load_library()
    assert lib.ddx_version() == 1
    p = ctypes.c_void_p()
    sh = (ctypes.c_int64 * 2)(8, 8)
    ln = (ctypes.c_double * 2)(1.0, 1.0)
    if _hip._device is None:
        rc = lib.ddx_plan_create(2, sh, ln, 0, ctypes.byref(p))
        assert rc == -5
        assert b"ddx_init" in lib.ddx_last_error()

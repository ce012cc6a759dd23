"""The C-ABI library loads, exports every entry point include/ehss_hip.h declares,
and reports errors through status codes (no GPU needed)."""
import ctypes as ct
import re
from pathlib import Path

import pytest

from ehss_amd import _lib as L

HEADER = Path(__file__).resolve().parents[1] / "include" / "ehss_hip.h"


def declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ehss_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared()
    assert "ehss_trace" in names and "ehss_solve" in names
    lib = L.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"


def test_abi_version_and_error_buffer():
    assert L.fn("ehss_abi_version")() == 1
    buf = ct.create_string_buffer(8)
    assert L.fn("ehss_last_error")(buf, len(buf)) == 0
    assert L.fn("ehss_last_error")(None, 0) == L.EHSS_ERR_INVALID


def test_invalid_arguments_fail_before_any_device_work():
    """Argument checks return EHSS_ERR_INVALID (or the device error when no GPU
    is present) with a message, never a crash."""
    null = ct.c_void_p()
    st = L.fn("ehss_scene_create")(0, null, null, 4, ct.byref(ct.c_void_p()))
    assert st in (L.EHSS_ERR_INVALID, L.EHSS_ERR_HIP)
    assert L.last_error()
    assert L.fn("ehss_scene_get_info")(None, None) == L.EHSS_ERR_INVALID
    assert "null" in L.last_error()


def test_no_cpu_fallback_without_device():
    """Without a GPU, compute entry points fail loudly instead of computing elsewhere."""
    try:
        L.device_info()
    except L.EHSSError as exc:
        assert exc.status == L.EHSS_ERR_HIP
    else:
        pytest.skip("a GPU is present")

"""CPU checks of the C-ABI library: it loads, exports every entry point that
include/ppvm_hip.h declares, validates arguments before touching a device, and
reports the absence of a device as an error (no CPU fallback)."""

import ctypes
import os
import re

import numpy as np
import pytest

from ppvm_amd import _native
from ppvm_amd import ops as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ppvm\w*)\s*\(", src)))


def test_header_declares_boundary():
    names = _declared("ppvm_hip.h")
    for must in ["ppvm_create", "ppvm_destroy", "ppvm_add_terms", "ppvm_rx", "ppvm_rzz", "ppvm_h",
                 "ppvm_cnot", "ppvm_truncate", "ppvm_overlap_with_zero", "ppvm_export_terms"]:
        assert must in names


@pytest.mark.parametrize("header,lib", [("ppvm_hip.h", "libppvm_hip.so")])
def test_library_exports_every_declared_symbol(header, lib):
    path = os.path.join(ROOT, "ppvm_amd", lib)
    assert os.path.exists(path), f"{path} not built"
    L = ctypes.CDLL(path)
    missing = [n for n in _declared(header) if not hasattr(L, n)]
    assert not missing, missing


def test_op_record_layout_matches_header():
    assert O.OP_DTYPE.itemsize == 144
    assert O.OP_DTYPE.fields["theta"][1] == 16
    assert O.OP_DTYPE.fields["p"][1] == 24


def _has_device():
    L = _native.lib()
    buf = ctypes.create_string_buffer(128)
    return L.ppvm_device_info(buf, 128) == 0


def test_no_device_is_an_error_not_a_fallback():
    if _has_device():
        pytest.skip("a GPU is present")
    L = _native.lib()
    h = ctypes.c_void_p()
    rc = L.ppvm_create(20, 1e-4, (1 << 64) - 1, 0, ctypes.byref(h))
    assert rc == _native.PPVM_ENODEV
    assert b"device" in L.ppvm_last_error()


def test_argument_validation_before_device():
    L = _native.lib()
    assert L.ppvm_create(0, 1e-4, 0, 0, ctypes.byref(ctypes.c_void_p())) == _native.PPVM_EINVAL
    assert L.ppvm_create(600, 1e-4, 0, 0, ctypes.byref(ctypes.c_void_p())) == _native.PPVM_EINVAL
    assert L.ppvm_create(10, -1.0, 0, 0, ctypes.byref(ctypes.c_void_p())) == _native.PPVM_EINVAL
    assert L.ppvm_rx(None, None, 0, 0.1) == _native.PPVM_EINVAL
    with pytest.raises(ValueError):
        _native.check(L.ppvm_truncate(None))


def test_lossy_entry_points_without_device():
    L = _native.lib()
    h = ctypes.c_void_p()
    mx = (1 << 64) - 1
    assert L.ppvm_lossy_create(0, 1e-4, mx, mx, 0, ctypes.byref(h)) == _native.PPVM_EINVAL
    assert L.ppvm_lossy_create(600, 1e-4, mx, mx, 0, ctypes.byref(h)) == _native.PPVM_EINVAL
    assert L.ppvm_lossy_create(10, -1.0, mx, mx, 0, ctypes.byref(h)) == _native.PPVM_EINVAL
    assert L.ppvm_lossy_apply(None, None, 0) == _native.PPVM_EINVAL
    if not _has_device():
        assert L.ppvm_lossy_create(20, 1e-4, mx, 2, 0, ctypes.byref(h)) == _native.PPVM_ENODEV
        assert b"device" in L.ppvm_lossy_last_error()


def test_lossy_words_layout_roundtrip():
    strings = ["XLZI" * 20, "I" * 79 + "L", "L" * 80, "_".join("IXYZ" * 20)]
    x, z, l = O.pack_lossy_strings(strings, 80)
    assert x.shape == (4, 2) and not np.any(l & (x | z))
    assert O.unpack_lossy_words(x, z, l, 80) == [s.replace("_", "") for s in strings]
    with pytest.raises(ValueError):
        O.pack_lossy_strings(["LQ"], 2)


def test_words_layout_roundtrip():
    strings = ["XYZI" * 20, "I" * 79 + "Y", "Z" * 80]
    x, z = O.pack_strings(strings, 80)
    assert x.shape == (3, 2)
    assert O.unpack_words(x, z, 80) == strings
    assert list(O.weight_of(x, z)) == [60, 1, 80]
    # qubit 64 lives in bit 0 of word 1
    x, z = O.pack_strings(["I" * 64 + "X" + "I" * 15], 80)
    assert int(x[0, 1]) == 1 and int(x[0, 0]) == 0
    assert O.storage_tier(20) == 2 and O.storage_tier(100) == 4 and O.storage_tier(127) == 5
    np.testing.assert_array_equal(z, 0)

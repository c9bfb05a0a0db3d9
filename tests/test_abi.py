"""The C-ABI library loads and exports its declared entry points (CPU only;
no call here touches a GPU)."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import sufkit_amd

HEADER = Path(__file__).resolve().parent.parent / "include" / "sufkit_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sfk_hip_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("sfk_hip_build", "sfk_hip_build_device", "sfk_hip_builder_create",
                 "sfk_hip_builder_destroy", "sfk_hip_abi_version"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = sufkit_amd.load_hip()
    for name in declared_functions():
        assert hasattr(lib, name), f"{name} declared in sufkit_hip.h but not exported"


def test_abi_version_and_device_count():
    lib = sufkit_amd.load_hip()
    assert lib.sfk_hip_abi_version() == 6
    assert lib.sfk_hip_device_count() >= 0


def _status(**kw):
    lib = sufkit_amd.load_hip()
    text = np.array([2, 3, 0], np.uint8)
    sa = np.zeros(3, np.uint64)
    err = ctypes.create_string_buffer(256)
    args = dict(n=3, bits=32, k=1, flags=3, sa_bits=64, isa=None, isa_bits=64, lcp=None, lcp_bits=64)
    args.update(kw)
    status = lib.sfk_hip_build(text.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), args["n"],
                               args["bits"], args["k"], args["flags"], sa.ctypes.data, args["sa_bits"],
                               args["isa"], args["isa_bits"], args["lcp"], args["lcp_bits"],
                               None, err, len(err))
    return status, err.value.decode()


@pytest.mark.parametrize("kw,needle", [
    (dict(bits=16), "construction width"),
    (dict(n=0), "empty"),
    (dict(sa_bits=40), "SA output width"),
    (dict(k=0), "sampling rate"),
])
def test_argument_validation_before_any_device_work(kw, needle):
    status, message = _status(**kw)
    assert status == 1  # ErrorCode::kInvalidInput + 1
    assert needle in message


def test_too_large_for_32bit_construction():
    status, message = _status(n=(1 << 32) + 5)
    assert status == 1 and "too large" in message


@pytest.mark.parametrize("n", [(1 << 32) + 5, (1 << 32) - (1 << 16) + 1, 1 << 40])
def test_longer_than_device_limit_is_unsupported(n):
    # 64-bit construction beyond SFK_HIP_MAX_TEXT_SYMBOLS: rejected before
    # any device work (one-element-per-work-item launches stay within the
    # 32-bit dispatch grid).
    status, message = _status(n=n, bits=64)
    assert status == 3 and "4294901760" in message  # kUnsupportedBackend + 1


def test_workspace_estimate_scales():
    lib = sufkit_amd.load_hip()
    small = lib.sfk_hip_workspace_bytes(1 << 20, 32)
    big = lib.sfk_hip_workspace_bytes(1 << 24, 32)
    wide = lib.sfk_hip_workspace_bytes(1 << 24, 64)
    assert small < big < wide


def test_no_fallback_when_library_missing(monkeypatch, tmp_path):
    monkeypatch.setattr(sufkit_amd, "_HIP", None)
    monkeypatch.setattr(sufkit_amd, "LIB_DIR", tmp_path)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        sufkit_amd.load_hip()

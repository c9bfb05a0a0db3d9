"""The allocation-free build (sfk_hip_build_device_noalloc) against the
allocating one (sfk_hip_build_device): identical SA/ISA/LCP on two calls in
a row, the second on a smaller text, so the shared workspace array holds the
first call's contents when the second starts."""
import ctypes

import numpy as np
import pytest

import sufkit_amd
from tests import synth

pytestmark = pytest.mark.gpu

FLAGS = sufkit_amd.SFK_HIP_RETAIN_ISA | sufkit_amd.SFK_HIP_RETAIN_LCP


def _lib():
    lib = sufkit_amd.load_hip()
    vp = ctypes.c_void_p
    lib.sfk_hip_build_device_noalloc.argtypes = [
        ctypes.c_int, vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, vp, ctypes.c_uint32,
        vp, ctypes.c_uint32, vp, ctypes.c_uint32, vp, ctypes.POINTER(sufkit_amd.HipStats), ctypes.c_char_p,
        ctypes.c_size_t]
    lib.sfk_hip_build_device_noalloc.restype = ctypes.c_int
    lib.sfk_hip_noalloc_workspace_bytes.argtypes = [ctypes.c_uint64, ctypes.c_int]
    lib.sfk_hip_noalloc_workspace_bytes.restype = ctypes.c_uint64
    return lib


def _noalloc(lib, d_text, n, outs, with_isa=True, with_lcp=True, bits=32, k=1, flags=FLAGS):
    err = ctypes.create_string_buffer(512)
    stats = sufkit_amd.HipStats()
    status = lib.sfk_hip_build_device_noalloc(
        0, d_text, n, bits, k, flags, outs[0], 32, outs[1] if with_isa else None, 32,
        outs[2] if with_lcp else None, 32, None, ctypes.byref(stats), err, len(err))
    return status, err.value.decode(errors="replace")


def _texts():
    big = synth.repeat_text(71, 980_000, cover=0.6, n_runs=6, contigs=5)
    small = synth.run_text(72, [("rand", 300), ("A", 5000), ("sep", 1), ("tandem3", 3000), ("N", 2000),
                                ("rand", 80_000)])
    return big, small


def test_noalloc_equals_allocating_build_twice_in_a_row(hip):
    lib = _lib()
    big, small = _texts()
    assert big.size <= 1_000_000 and small.size < big.size
    assert lib.sfk_hip_noalloc_workspace_bytes(1_000_000, 1) <= (64 << 20)
    assert lib.sfk_hip_noalloc_workspace_bytes(1_000_000, 0) <= (64 << 20)
    cap = big.size
    d_text = sufkit_amd.DeviceArray(cap + 64)
    got = [sufkit_amd.DeviceArray(cap * 4) for _ in range(3)]
    want = [sufkit_amd.DeviceArray(cap * 4) for _ in range(3)]
    builder = sufkit_amd.DeviceBuilder(0)
    try:
        for text in (big, small):
            n = int(text.size)
            d_text.upload(text)
            status, msg = _noalloc(lib, d_text.ptr, n, [a.ptr for a in got])
            assert status == 0, msg
            builder.build(d_text.ptr, n, 32, want[0].ptr, 32, want[1].ptr, 32, want[2].ptr, 32, FLAGS)
            for g, w in zip(got, want):
                assert np.array_equal(g.download(np.uint32, n), w.download(np.uint32, n))
        # SA only, then SA + ISA without LCP, on the small text again
        n = int(small.size)
        status, msg = _noalloc(lib, d_text.ptr, n, [a.ptr for a in got], with_isa=False, with_lcp=False,
                               flags=0)
        assert status == 0, msg
        assert np.array_equal(got[0].download(np.uint32, n), want[0].download(np.uint32, n))
        status, msg = _noalloc(lib, d_text.ptr, n, [a.ptr for a in got], with_lcp=False)
        assert status == 0, msg
        assert np.array_equal(got[1].download(np.uint32, n), want[1].download(np.uint32, n))
    finally:
        builder.close()


def test_noalloc_rejects_what_it_cannot_hold(hip):
    lib = _lib()
    d = sufkit_amd.DeviceArray(64)
    outs = [d.ptr] * 3
    status, msg = _noalloc(lib, d.ptr, 1_000_001, outs)
    assert status == 1 and "at most 1000000" in msg
    status, msg = _noalloc(lib, d.ptr, 1000, outs, bits=64)
    assert status == 1
    status, msg = _noalloc(lib, d.ptr, 1000, outs, k=2)
    assert status == 1

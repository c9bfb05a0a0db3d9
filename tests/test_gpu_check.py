"""The device digest and the device index checker (sfk_hip_digest,
sfk_hip_check_index) against the oracle, and the failure paths they map to
the reference's ErrorCode::kBuildFailure (a suffix array that is not a
permutation with an inverse, suffix_array.cpp:852-881).

The arrays fed to the checker are the oracle's SA/ISA/LCP, made on the CPU,
some of them altered on the CPU before the upload; the checker must accept
exactly the unaltered ones."""
import numpy as np
import pytest

import oracle
import sufkit_amd
from tests import synth

pytestmark = pytest.mark.gpu

BUILD_FAILURE = 6  # ErrorCode::kBuildFailure + 1


@pytest.mark.parametrize("bits,count", [(8, 0), (8, 1000), (32, 1), (32, 123457), (64, 99991)])
def test_device_digest_equals_oracle(hip, bits, count):
    rng = np.random.default_rng(count + bits)
    dtype = {8: np.uint8, 32: np.uint32, 64: np.uint64}[bits]
    values = rng.integers(0, np.iinfo(dtype).max, size=count, dtype=dtype, endpoint=True)
    d = sufkit_amd.DeviceArray(max(values.nbytes, 8))
    if count:
        d.upload(values)
    assert sufkit_amd.device_digest(d.ptr, count, bits) == oracle.sum_digest(values.astype(np.uint64))


def test_digest_sees_one_changed_and_two_swapped_values(hip):
    values = np.arange(1, 50001, dtype=np.uint64)
    base = oracle.sum_digest(values)
    changed = values.copy()
    changed[777] += 1
    swapped = values.copy()
    swapped[[10, 20000]] = swapped[[20000, 10]]
    d = sufkit_amd.DeviceArray(values.nbytes)
    digests = set()
    for arr in (values, changed, swapped):
        d.upload(arr)
        digests.add(sufkit_amd.device_digest(d.ptr, arr.size, 64))
    assert len(digests) == 3 and base in digests
    # equal values in another width give the same digest
    d32 = sufkit_amd.DeviceArray(values.size * 4)
    d32.upload(values.astype(np.uint32))
    assert sufkit_amd.device_digest(d32.ptr, values.size, 32) == base


def _upload(arr):
    d = sufkit_amd.DeviceArray(max(arr.nbytes, 8))
    d.upload(arr)
    return d


def _check(text, sa, isa, lcp, bits=32):
    dt = {32: np.uint32, 64: np.uint64}[bits]
    bufs = [_upload(text)] + [_upload(np.ascontiguousarray(a, dtype=dt)) for a in (sa, isa)]
    d_lcp = _upload(np.ascontiguousarray(lcp, dtype=dt)) if lcp is not None else None
    rep = sufkit_amd.check_index(bufs[0].ptr, text.size, bufs[1].ptr, bits, bufs[2].ptr, bits,
                                 d_lcp.ptr if d_lcp else None, bits, raise_on_failure=False)
    return rep


TEXTS = {
    "fixture": None,
    "random": lambda: synth.random_text(5, 20000),
    "symbols": lambda: synth.random_symbols(6, 5000),
    "runs": lambda: synth.run_text(7, [("rand", 40), ("A", 3000), ("sep", 1), ("tandem3", 900), ("N", 700)]),
    "repeats": lambda: synth.repeat_text(8, 30000, cover=0.7, n_runs=4, contigs=3),
}


@pytest.fixture(scope="module", params=sorted(TEXTS))
def built(request, legacy_fixtures):
    if request.param == "fixture":
        text = np.array(legacy_fixtures["legacy-sa-v1.1"]["vectors"]["text"], np.uint8)
    else:
        text = TEXTS[request.param]()
    sa, isa, lcp = oracle.fast_build(text)
    return text, sa, isa, lcp


@pytest.mark.parametrize("bits", [32, 64])
def test_correct_index_passes(hip, built, bits):
    text, sa, isa, lcp = built
    rep = _check(text, sa, isa, lcp, bits)
    assert rep["status"] == 0 and rep["rows"] == text.size and rep["first_failed_row"] == text.size, rep
    rep = _check(text, sa, isa, None, bits)
    assert rep["status"] == 0 and rep["lcp_checked"] == 0


def test_swapped_rows_with_consistent_inverse_fail_order(hip, built):
    text, sa, isa, lcp = built
    sa2, isa2 = sa.copy(), isa.copy()
    r = text.size // 2
    sa2[[r, r + 1]] = sa2[[r + 1, r]]
    isa2[sa2[r]], isa2[sa2[r + 1]] = r, r + 1
    rep = _check(text, sa2, isa2, None)
    assert rep["status"] == BUILD_FAILURE and rep["order_failures"] >= 1 and rep["inverse_failures"] == 0
    assert "out of order" in rep["message"]


def test_swapped_rows_alone_fail_inverse(hip, built):
    text, sa, isa, lcp = built
    sa2 = sa.copy()
    sa2[[3, text.size - 2]] = sa2[[text.size - 2, 3]]
    rep = _check(text, sa2, isa, lcp)
    assert rep["status"] == BUILD_FAILURE and rep["inverse_failures"] == 2
    assert rep["first_failed_row"] == 3


def test_out_of_range_values_fail_without_faulting(hip, built):
    text, sa, isa, lcp = built
    sa2, isa2 = sa.copy(), isa.copy()
    sa2[5] = text.size + 1000
    isa2[7] = 2 ** 32 - 1
    rep = _check(text, sa2, isa2, lcp)
    assert rep["status"] == BUILD_FAILURE and rep["range_failures"] >= 2


@pytest.mark.parametrize("delta", [1, -1, 1000, 2 ** 31])
def test_wrong_lcp_values_fail(hip, built, delta):
    text, sa, isa, lcp = built
    lcp2 = lcp.astype(np.int64)
    rows = [r for r in range(1, text.size) if lcp2[r] + delta >= 0][:: max(1, text.size // 7)]
    for r in rows:
        lcp2[r] += delta
    rep = _check(text, sa, isa, lcp2.astype(np.uint64))
    assert rep["status"] == BUILD_FAILURE
    assert rep["lcp_failures"] + rep["lcp_prefix_failures"] == len(rows)


def test_one_too_large_is_seen_by_the_prefix_hash(hip, built):
    # Rows whose suffixes also differ one symbol after their mismatch: with
    # LCP + 1 the mismatch test passes, and only the hash of the first
    # LCP + 1 symbols (which holds the real mismatch) can fail them.
    text, sa, isa, lcp = built
    n = text.size
    rows = []
    for r in range(1, n):
        a, b, l = int(sa[r - 1]), int(sa[r]), int(lcp[r])
        if l + 1 < n - max(a, b) and text[a + l + 1] != text[b + l + 1]:
            rows.append(r)
    rows = rows[:: max(1, len(rows) // 9)]
    assert rows
    lcp2 = lcp.copy()
    lcp2[rows] += 1
    rep = _check(text, sa, isa, lcp2)
    assert rep["status"] == BUILD_FAILURE
    assert rep["lcp_prefix_failures"] == len(rows) and rep["lcp_failures"] == 0


def test_nonzero_first_lcp_fails(hip, built):
    text, sa, isa, lcp = built
    lcp2 = lcp.copy()
    lcp2[0] = 1
    rep = _check(text, sa, isa, lcp2)
    assert rep["status"] == BUILD_FAILURE and rep["lcp_failures"] == 1 and rep["first_failed_row"] == 0


def test_invalid_arguments(hip):
    with pytest.raises(sufkit_amd.SufkitError) as e:
        sufkit_amd.check_index(0, 10, 0, 32, 0, 32)
    assert e.value.status == 1


def test_allocation_failure_is_a_build_failure_and_recoverable(hip):
    # An out-of-memory device allocation in the middle of a build surfaces as
    # ErrorCode::kBuildFailure (as a failed CaPS allocation does,
    # caps_backend.cpp:68-76); the same builder then builds correctly.
    import ctypes
    from sufkit_amd import api

    lib = sufkit_amd.load_hip()
    lib.sfk_hip_inject_allocation_failure.argtypes = [ctypes.c_longlong]
    lib.sfk_hip_inject_allocation_failure.restype = None
    text = synth.repeat_text(41, 50000, cover=0.5, n_runs=3, contigs=2)
    n = text.size
    d_text = sufkit_amd.DeviceArray(n + 64)
    d_text.upload(text)
    outs = [sufkit_amd.DeviceArray(n * 4) for _ in range(3)]
    osa, oisa, olcp = oracle.fast_build(text)
    with sufkit_amd.DeviceBuilder(0) as b:
        for after in (0, 3, 7):
            lib.sfk_hip_inject_allocation_failure(after)
            with pytest.raises(sufkit_amd.SufkitError) as e:
                b.build(d_text.ptr, n, 32, outs[0].ptr, 32, outs[1].ptr, 32, outs[2].ptr, 32)
            assert e.value.status == BUILD_FAILURE and "allocation" in str(e.value)
        lib.sfk_hip_inject_allocation_failure(-1)
        b.build(d_text.ptr, n, 32, outs[0].ptr, 32, outs[1].ptr, 32, outs[2].ptr, 32)
    assert np.array_equal(outs[0].download(np.uint32, n).astype(np.uint64), osa)
    assert np.array_equal(outs[2].download(np.uint32, n).astype(np.uint64), olcp)
    # the host Build maps it to ErrorCode::kBuildFailure
    ref = api.GenomeReference.from_encoded(text, synth.contig_lengths(text))
    lib.sfk_hip_inject_allocation_failure(1)
    try:
        with pytest.raises(sufkit_amd.SufkitError) as e:
            api.SuffixArray.build(ref, api.BuildOptions(backend=api.BACKEND_HIP))
        assert e.value.code == "build_failure"
    finally:
        lib.sfk_hip_inject_allocation_failure(-1)


@pytest.mark.parametrize("bits,flags,k,path", [
    (32, 3, 1, "marked-row LCP (pair-multiset hash)"),
    (64, 3, 1, "marked-row LCP, 64-bit construction"),
    (32, 1, 1, "no LCP (ValidateKernel)"),
    (32, 3, 2, "sampled K=2 (full PLCP pass)"),
])
def test_swapped_sa_rows_fail_the_build(hip, bits, flags, k, path):
    # Two SA rows swapped after the doubling (the ISA left as built): every
    # build path's own permutation check must refuse it with kBuildFailure,
    # as RepackSuffixArray's check does (suffix_array.cpp:852-881); the next
    # build, unarmed, is exact.
    import ctypes
    lib = sufkit_amd.load_hip()
    lib.sfk_hip_inject_sa_swap.argtypes = [ctypes.c_int]
    lib.sfk_hip_inject_sa_swap.restype = None
    text = synth.repeat_text(43, 40000, cover=0.6, n_runs=4, contigs=3)
    n = text.size
    rows = (n + k - 1) // k
    d_text = sufkit_amd.DeviceArray(n + 64)
    d_text.upload(text)
    outs = [sufkit_amd.DeviceArray(rows * 4) for _ in range(3)]
    lcp = outs[2].ptr if flags & 2 else None
    with sufkit_amd.DeviceBuilder(0) as b:
        lib.sfk_hip_inject_sa_swap(1)
        try:
            with pytest.raises(sufkit_amd.SufkitError) as e:
                b.build(d_text.ptr, n, bits, outs[0].ptr, 32, outs[1].ptr, 32, lcp, 32, flags, k)
        finally:
            lib.sfk_hip_inject_sa_swap(0)
        assert e.value.status == BUILD_FAILURE and "permutation" in str(e.value), path
        b.build(d_text.ptr, n, bits, outs[0].ptr, 32, outs[1].ptr, 32, lcp, 32, flags, k)
    osa, oisa, olcp = oracle.fast_build(text)
    if k == 1:
        assert np.array_equal(outs[0].download(np.uint32, n).astype(np.uint64), osa)
        if flags & 2:
            assert np.array_equal(outs[2].download(np.uint32, n).astype(np.uint64), olcp)

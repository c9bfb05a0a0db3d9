"""Serving loop: continuous batching of build requests on one GPU, resident
indexes, device count queries.  Each request's SA/ISA/LCP must be identical
whether it is built alone or in the mix, and every count must equal a naive
count over the text."""
import numpy as np
import pytest

import sufkit_amd
from tests import synth

pytestmark = pytest.mark.gpu


def fnv(arr: np.ndarray) -> int:
    h = 1469598103934665603
    for b in arr.tobytes():
        h = ((h ^ b) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def render(text: np.ndarray) -> str:
    return "".join("$|ACGTN"[c] for c in text.tolist())


def naive_count(s: str, pattern: str) -> int:
    count, start = 0, s.find(pattern)
    while start != -1:
        count += 1
        start = s.find(pattern, start + 1)
    return count


def requests(seed=4, count=6):
    rng = np.random.default_rng(seed)
    texts = [synth.repeat_text(seed * 10 + i, int(rng.integers(3000, 20000)), cover=0.5, n_runs=3, contigs=2)
             for i in range(count)]
    arrivals = np.cumsum(rng.exponential(0.01, size=count)).tolist()
    return texts, arrivals


def test_results_identical_alone_and_in_the_mix(hip):
    texts, arrivals = requests()
    mix, msum, msamples = hip.serve(texts, arrivals, slots=3, queries_per_index=4000, sample_queries=150,
                                    digest_outputs=True)
    alone, asum, _ = hip.serve(texts, arrivals, slots=1, queries_per_index=4000, sample_queries=0,
                               digest_outputs=True)
    assert all(r["status"] == 0 for r in mix + alone)
    for text, m, a in zip(texts, mix, alone):
        assert m["output_hash"] == a["output_hash"]
        assert m["count_hash"] == a["count_hash"] and m["count_sum"] == a["count_sum"]
        sa, isa, lcp, _ = hip.hip_build(text)
        assert m["output_hash"] == [fnv(sa), fnv(isa), fnv(lcp)]
    assert msum["latency_p99"] >= msum["latency_p50"] > 0
    assert msum["requests_failed"] == 0
    # sampled counts vs a naive scan of the text
    patterns, lengths, counts = msamples
    for text, pats, lens, cnts in zip(texts, patterns, lengths, counts):
        s = render(text)
        for code, length, got in zip(pats, lens, cnts):
            assert got == naive_count(s, hip.decode_pattern(code, length))


def test_tiny_budget_admits_one_at_a_time(hip):
    texts, arrivals = requests(seed=9, count=4)
    res, summary, _ = hip.serve(texts, [0.0] * 4, slots=4, hbm_budget_bytes=1, queries_per_index=100,
                                keep_resident=False)
    assert all(r["status"] == 0 for r in res)
    assert summary["max_concurrent"] == 1


def test_small_budget_evicts_resident_indexes(hip):
    texts, arrivals = requests(seed=11, count=5)
    one = max(t.size for t in texts) * 40
    res, summary, _ = hip.serve(texts, arrivals, slots=2, hbm_budget_bytes=one, queries_per_index=100,
                                keep_resident=True)
    assert all(r["status"] == 0 for r in res)
    assert summary["evicted"] >= 1


def test_count_queries_on_resident_index(hip):
    text = synth.run_text(3, [("rand", 3000), ("A", 500), ("sep", 1), ("N", 200), ("tandem3", 600),
                              ("rand", 1000)])
    n = text.size
    d_text = hip.DeviceArray(n + 64)
    d_text.upload(text)
    d_sa, d_isa, d_lcp = hip.DeviceArray(4 * n), hip.DeviceArray(4 * n), hip.DeviceArray(4 * n)
    builder = hip.DeviceBuilder(0)
    builder.build(d_text.ptr, n, 32, d_sa.ptr, 32, d_isa.ptr, 32, d_lcp.ptr, 32)
    builder.close()
    import ctypes
    lib = hip.load_hip()
    lib.sfk_hip_index_create.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                                         ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p), ctypes.c_char_p,
                                         ctypes.c_size_t]
    lib.sfk_hip_index_count.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                        ctypes.c_void_p, ctypes.c_void_p]
    lib.sfk_hip_index_destroy.argtypes = [ctypes.c_void_p]
    index = ctypes.c_void_p()
    err = ctypes.create_string_buffer(256)
    assert lib.sfk_hip_index_create(0, d_text.ptr, n, d_sa.ptr, 32, ctypes.byref(index), err, 256) == 0
    s = render(text)
    rng = np.random.default_rng(1)
    pats, codes, lens = [], [], []
    for k in range(400):
        length = int(rng.integers(1, 33))
        if k % 2 == 0:
            start = int(rng.integers(0, n - length - 1))
            p = "".join(ch if ch in "ACGT" else "A" for ch in s[start:start + length])
        else:
            p = "".join("ACGT"[i] for i in rng.integers(0, 4, size=length))
        code = 0
        for j, ch in enumerate(p):
            code |= "ACGT".index(ch) << (62 - 2 * j)
        pats.append(p)
        codes.append(code)
        lens.append(length)
    d_p = hip.DeviceArray(8 * len(pats))
    d_p.upload(np.array(codes, np.uint64))
    d_l = hip.DeviceArray(len(pats))
    d_l.upload(np.array(lens, np.uint8))
    d_c = hip.DeviceArray(4 * len(pats))
    assert lib.sfk_hip_index_count(index, d_p.ptr, d_l.ptr, len(pats), d_c.ptr, None) == 0
    got = d_c.download(np.uint32, len(pats))
    lib.sfk_hip_index_destroy(index)
    for p, g in zip(pats, got):
        assert g == naive_count(s, p), p

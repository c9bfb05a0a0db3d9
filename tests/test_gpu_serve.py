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


def test_fm_and_sa_count_queries_agree(hip):
    # The served loop answers by FM backward search by default; binary
    # search over the resident SA gives the same counts for every query.
    texts, arrivals = requests(seed=6, count=4)
    fm, fsum, _ = hip.serve(texts, arrivals, slots=2, queries_per_index=20000, fm_queries=True)
    sa, ssum, _ = hip.serve(texts, arrivals, slots=2, queries_per_index=20000, fm_queries=False)
    assert all(r["status"] == 0 for r in fm + sa)
    for f, s in zip(fm, sa):
        assert f["count_hash"] == s["count_hash"] and f["count_sum"] == s["count_sum"] > 0


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


SERVED_SMALLEST = [6, 15, 1, 7]  # the four smallest requests of the served workload (1.0-4.6 M symbols)


def test_served_workload_requests_match_the_oracle(hip):
    import oracle

    for r in SERVED_SMALLEST[:3]:
        text, _ = hip.synthetic_text(5, 20260822, r, 1.0)
        sa, isa, lcp, _ = hip.hip_build(text)
        osa, oisa, olcp = oracle.fast_build(text)
        assert np.array_equal(sa.astype(np.uint64), osa), f"request {r}: SA differs"
        assert np.array_equal(isa.astype(np.uint64), oisa), f"request {r}: ISA differs"
        assert np.array_equal(lcp.astype(np.uint64), olcp), f"request {r}: LCP differs"


def test_served_workload_requests_identical_in_the_mix(hip):
    # The smallest served requests built concurrently on four slots, one
    # after another on one reused builder, and each on a fresh builder give
    # the same SA, ISA, LCP and counts.
    texts = [hip.synthetic_text(5, 20260822, r, 1.0)[0] for r in SERVED_SMALLEST]
    arrivals = [0.0, 0.001, 0.002, 0.003]
    q = 20000
    mix, msum, ms = hip.serve(texts, arrivals, slots=4, queries_per_index=q, digest_outputs=True, sample_queries=q)
    alone, _, als = hip.serve(texts, arrivals, slots=1, queries_per_index=q, digest_outputs=True, sample_queries=q)
    # A lone request is request 0 of its run: its query seed is the one the
    # loop derives for request i (seed ^ i * 0x9E3779B97F4A7C15).
    fresh = [hip.serve([t], [0.0], slots=1, queries_per_index=q, digest_outputs=True, sample_queries=q,
                       query_seed=7 ^ ((i * 0x9E3779B97F4A7C15) & ((1 << 64) - 1)))
             for i, t in enumerate(texts)]
    assert all(r["status"] == 0 for r in mix + alone + [f[0][0] for f in fresh]) and msum["max_concurrent"] >= 2
    for i, (text, m, a, f) in enumerate(zip(texts, mix, alone, fresh)):
        fr = f[0][0]
        if m["count_hash"] != fr["count_hash"] or a["count_hash"] != fr["count_hash"]:
            s = render(text)
            for name, samples in (("mix", ms), ("alone", als)):
                diff = np.nonzero(samples[2][i] != f[2][2][0])[0]
                for k in diff[:4].tolist():
                    p = hip.decode_pattern(int(samples[0][i][k]), int(samples[1][i][k]))
                    print(f"request {i} ({name}) query {k} {p}: {samples[2][i][k]} vs fresh {f[2][2][0][k]}, "
                          f"scan {naive_count(s, p)}; patterns equal {samples[0][i][k] == f[2][0][0][k]}",
                          flush=True)
        for key in ("output_hash", "count_hash", "count_sum"):
            assert m[key] == fr[key] and a[key] == fr[key], (text.size, key)


@pytest.mark.parametrize("fm", [False, True])
def test_runs_of_one_base_are_counted(hip, fm):
    # Poly-A and period-2 blocks (the adversarial block of the served
    # requests) hold long runs: counts of A^L and (AT)^L against a scan.
    text, _ = hip.synthetic_text(5, 20260822, 15, 1.0)
    s = render(text)
    sa, _, _, _ = hip.hip_build(text)
    pats = ["A" * L for L in (1, 5, 8, 9, 10, 11, 20, 28, 32)] + ["AT" * L for L in (1, 5, 14, 16)] + \
        ["T" * 28, "C" * 9, "ACGTACGT"]
    got = hip.index_count(text, sa, pats, fm=fm)
    for p, g in zip(pats, got):
        assert g == naive_count(s, p), (p, g)


def test_strict_budget_refuses_an_oversized_request(hip):
    # Under strict_budget a request whose footprint alone exceeds the HBM
    # budget is refused with ErrorCode::kInvalidInput (status 1) and a
    # message naming the budget; the other requests are served.
    small = [synth.random_text(80 + i, 20_000) for i in range(3)]
    big = synth.random_text(90, 3_000_000)
    texts = [small[0], big, small[1], small[2]]
    budget = 64 << 20  # holds a 20 k request, not a 3 M one (about 100 bytes per symbol)
    res, summ, _ = hip.serve(texts, [0.0, 0.0005, 0.001, 0.0015], slots=2, queries_per_index=1000,
                             hbm_budget_bytes=budget, strict_budget=True)
    assert [r["status"] for r in res] == [0, 1, 0, 0]
    assert summ["requests_failed"] == 1
    msg = summ["first_error"].decode()
    assert msg.startswith("request 1 (status 1)") and "exceeds the HBM budget" in msg

"""FM-index from the GPU SA: FmIndex queries (EqualRange, Count, CountBatch,
Locate) against the reference's own FmIndex (golden vectors from the
compiled reference, tests/golden/fm_golden.json) and against the oracle's
exact search by definition.

CPU tests pin the oracle's search functions: against the reference's
FmIndex where oracle/_ref is built, against the golden vectors, and against
the SA interval of the brute-force SA.  GPU tests run the device FM-index
through the host library's C ABI (and the device index's C ABI directly)."""
from __future__ import annotations

import ctypes
import json

import numpy as np
import pytest

import oracle
from tests import synth
from tests.conftest import GOLDEN

FM_GOLDEN = json.loads((GOLDEN / "fm_golden.json").read_text())


def case_text(case) -> tuple[np.ndarray, np.ndarray]:
    seqs = [seq.encode() for _, seq in case["records"]]
    text, _ = oracle.encode_records(seqs)
    starts = np.cumsum([0] + [len(s) + 1 for s in seqs[:-1]]).astype(np.uint64)
    return text, starts


def as_hits(result) -> dict:
    return {"total_hits": result["total_hits"], "hits": [tuple(h) for h in result["hits"]],
            "truncated": result["truncated"]}


# ------------------------------------------------------------ CPU: oracle pins

@pytest.mark.parametrize("case_index", range(len(FM_GOLDEN["cases"])))
def test_oracle_search_reproduces_the_golden_vectors(case_index):
    case = FM_GOLDEN["cases"][case_index]
    text, starts = case_text(case)
    for entry in case["patterns"]:
        p = entry["pattern"]
        assert list(oracle.equal_range(text, p)) == entry["equal_range"], p
        rc = oracle.reverse_complement(p)
        fwd, rev = oracle.occurrences(text, p).size, oracle.occurrences(text, rc).size
        both = fwd if rc.upper() == p.upper() else fwd + rev
        assert [fwd, rev, both] == entry["count"], p
        for key, expected in entry["locate"].items():
            strands, limit = key.split(":")
            limit = None if limit == "None" else int(limit)
            got = oracle.locate(text, starts, p, int(strands), limit)
            assert as_hits(got) == as_hits(expected), (p, key)


def test_oracle_equal_range_is_the_brute_force_sa_interval():
    rng = np.random.default_rng(7)
    text, _ = oracle.encode_records([synth_bases(rng, 700), synth_bases(rng, 300, n_runs=2), b"ACGTACGT"])
    sa = oracle.sa_naive(text)
    for length in (1, 2, 3, 6, 9):
        for _ in range(12):
            p = "".join("ACGT"[c] for c in rng.integers(0, 4, length))
            codes = oracle.encode_pattern(p)
            rows = [r for r in range(text.size)
                    if np.array_equal(text[sa[r]:sa[r] + length], codes)]
            expected = (rows[0], rows[-1] + 1) if rows else (0, 0)
            assert rows == list(range(expected[0], expected[1])) or not rows
            assert oracle.equal_range(text, p) == expected, p


def test_planted_mistake_in_the_oracle_is_caught():
    # A range shifted by one row, or hits left in global coordinates, fail
    # the golden comparison that pins the oracle.
    case = FM_GOLDEN["cases"][1]
    text, starts = case_text(case)
    found = [e for e in case["patterns"] if e["equal_range"] != [0, 0] and e["locate"].get("0:None")]
    assert found
    e = found[0]
    shifted = [x + 1 for x in oracle.equal_range(text, e["pattern"])]
    assert shifted != e["equal_range"]
    e = next(e for e in found if any(h[0] > 0 for h in e["locate"]["0:None"]["hits"]))
    wrong = oracle.locate(text, np.zeros_like(starts), e["pattern"], 0)
    assert as_hits(wrong) != as_hits(e["locate"]["0:None"])


@pytest.mark.skipif(not oracle.ref_fm_available(), reason="oracle/_ref reference library not built here")
def test_golden_vectors_match_a_fresh_reference_run():
    case = FM_GOLDEN["cases"][1]
    fm = oracle.RefFmIndex([(name, seq.encode()) for name, seq in case["records"]])
    assert fm.build_status == 0
    for entry in case["patterns"][::5]:
        assert list(fm.equal_range(entry["pattern"])) == entry["equal_range"]
        assert [fm.count(entry["pattern"], s) for s in range(3)] == entry["count"]


@pytest.mark.skipif(not oracle.ref_fm_available(), reason="oracle/_ref reference library not built here")
def test_oracle_matches_the_reference_fm_on_random_references():
    rng = np.random.default_rng(11)
    for trial in range(3):
        seqs = [synth_bases(rng, int(rng.integers(50, 3000)), n_runs=int(rng.integers(0, 3)))
                for _ in range(int(rng.integers(1, 5)))]
        records = [(f"c{i}", s) for i, s in enumerate(seqs)]
        text, _ = oracle.encode_records(seqs)
        starts = np.cumsum([0] + [len(s) + 1 for s in seqs[:-1]]).astype(np.uint64)
        fm = oracle.RefFmIndex(records)
        for _ in range(25):
            length = int(rng.integers(1, 14))
            p = "".join("ACGT"[c] for c in rng.integers(0, 4, length))
            assert fm.equal_range(p) == oracle.equal_range(text, p)
            s = int(rng.integers(0, 3))
            assert as_hits(fm.locate(p, s, None)) == as_hits(oracle.locate(text, starts, p, s)), (p, s)


def synth_bases(rng, size: int, n_runs: int = 0) -> bytes:
    seq = bytearray(b"ACGT"[c] for c in rng.integers(0, 4, size))
    if size > 40:
        unit = bytes(seq[:12])
        for at in rng.integers(0, size - 12, 3):
            seq[at:at + 12] = unit
    for _ in range(n_runs):
        at = int(rng.integers(0, max(1, size - 20)))
        run = min(int(rng.integers(1, 20)), size - at)
        seq[at:at + run] = b"N" * run
    return bytes(seq)


# ------------------------------------------------------------ GPU: device FM

def device_fm(records):
    from sufkit_amd import api

    ref = api.GenomeReference.from_records([(name, seq if isinstance(seq, bytes) else seq.encode())
                                            for name, seq in records])
    return api.FmIndex.build(ref)


@pytest.mark.gpu
@pytest.mark.parametrize("case_index", range(len(FM_GOLDEN["cases"])))
def test_device_fm_matches_the_reference_golden_vectors(hip, case_index):
    case = FM_GOLDEN["cases"][case_index]
    fm = device_fm(case["records"])
    patterns = [e["pattern"] for e in case["patterns"]]
    for entry in case["patterns"]:
        p = entry["pattern"]
        assert list(fm.equal_range(p)) == entry["equal_range"], p
        assert [fm.count(p, s) for s in range(3)] == entry["count"], p
        for key, expected in entry["locate"].items():
            strands, limit = key.split(":")
            limit = None if limit == "None" else int(limit)
            assert as_hits(fm.locate(p, int(strands), limit)) == as_hits(expected), (p, key)
    assert [list(r) for r in fm.equal_range_batch(patterns)] == [e["equal_range"] for e in case["patterns"]]
    for key, expected in case["count_batch"].items():
        strands, width = map(int, key.split(":"))
        assert fm.count_batch(patterns, strands, width) == expected, key


@pytest.mark.gpu
def test_device_fm_error_codes_match_the_reference(hip):
    from sufkit_amd import SufkitError, api

    case = FM_GOLDEN["cases"][0]
    errors = case["errors"]
    fm = device_fm(case["records"])

    def code(fn):
        with pytest.raises(SufkitError) as e:
            fn()
        return e.value.status

    assert code(lambda: fm.equal_range("")) == errors["equal_range_empty"]
    assert code(lambda: fm.equal_range("ACGN")) == errors["equal_range_non_acgt"]
    assert code(lambda: fm.count("ACGT", 3)) == errors["count_bad_strand"]
    assert code(lambda: fm.count_batch(["ACGT"], 0, 300)) == errors["count_batch_width"]
    assert code(lambda: fm.count_batch(["ACGT", "AXG"])) == errors["count_batch_bad_pattern"]
    assert code(lambda: fm.locate("ACGT", 3)) == errors["locate_bad_strand"]
    ref = api.GenomeReference.from_records([(n, s.encode()) for n, s in case["records"]])
    assert code(lambda: api.FmIndex.build(ref, api.FM_SDSL_CSA_SADA)) == errors["build_sada"]


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_device_fm_matches_the_oracle_on_random_references(hip, seed):
    rng = np.random.default_rng(seed)
    seqs = [synth_bases(rng, int(rng.integers(1, 5000)), n_runs=int(rng.integers(0, 4)))
            for _ in range(int(rng.integers(1, 6)))]
    text, _ = oracle.encode_records(seqs)
    starts = np.cumsum([0] + [len(s) + 1 for s in seqs[:-1]]).astype(np.uint64)
    fm = device_fm([(f"c{i}", s) for i, s in enumerate(seqs)])
    plain = text[(text >= 2) & (text <= 5)]
    patterns = []
    for length in (1, 2, 3, 7, 9, 10, 11, 13, 24, 48, 70):
        for _ in range(4):
            if plain.size > length:
                at = int(rng.integers(0, plain.size - length))
                patterns.append("".join("ACGT"[c - 2] for c in plain[at:at + length]))
            patterns.append("".join("ACGT"[c] for c in rng.integers(0, 4, length)))
    for p in patterns:
        assert fm.equal_range(p) == oracle.equal_range(text, p), p
        s = int(rng.integers(0, 3))
        assert as_hits(fm.locate(p, s, None)) == as_hits(oracle.locate(text, starts, p, s)), (p, s)
    counts = fm.count_batch(patterns, 2, 0)
    for p, c in zip(patterns, counts):
        rc = oracle.reverse_complement(p)
        fwd = oracle.occurrences(text, p).size
        assert c == (fwd if rc == p else fwd + oracle.occurrences(text, rc).size), p


@pytest.mark.gpu
def test_device_fm_on_a_run_heavy_reference_with_long_patterns(hip):
    # Long runs give long backward searches whose intervals stay wide; the
    # record boundaries (64 rows) are crossed at every step.
    seq = b"A" * 3000 + b"AT" * 700 + b"ACGTTGCA" * 200 + b"N" * 40 + b"A" * 513
    fm = device_fm([("runs", seq), ("tail", b"ACGT" * 17)])
    text, starts = oracle.encode_records([seq, b"ACGT" * 17])[0], np.array([0, len(seq) + 1], np.uint64)
    for p in ["A" * 1, "A" * 63, "A" * 64, "A" * 65, "A" * 513, "A" * 514, "AT" * 40, "TA" * 41,
              "ACGTTGCA" * 20, "ACGT" * 17, "ACGT" * 18, "ATACGT"]:
        assert fm.equal_range(p) == oracle.equal_range(text, p), p
        assert as_hits(fm.locate(p, 2, 5)) == as_hits(oracle.locate(text, starts, p, 2, 5)), p


def index_lib():
    import sufkit_amd

    lib = sufkit_amd.load_hip()
    vp = ctypes.c_void_p
    lib.sfk_hip_index_create.argtypes = [ctypes.c_int, vp, ctypes.c_uint64, vp, ctypes.c_uint32,
                                         ctypes.POINTER(vp), ctypes.c_char_p, ctypes.c_size_t]
    lib.sfk_hip_index_destroy.argtypes = [vp]
    lib.sfk_hip_index_destroy.restype = None
    lib.sfk_hip_index_build_fm.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
    lib.sfk_hip_index_fm_header.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64)]
    lib.sfk_hip_index_fm_ranges.argtypes = [vp, vp, vp, vp, ctypes.c_uint64, vp, vp]
    lib.sfk_hip_index_fm_count.argtypes = [vp, vp, vp, ctypes.c_uint64, vp, vp]
    lib.sfk_hip_index_count.argtypes = [vp, vp, vp, ctypes.c_uint64, vp, vp]
    lib.sfk_hip_index_generate_queries.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32,
                                                   ctypes.c_uint32, vp, vp, vp]
    lib.sfk_hip_index_sa_rows.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
    return lib


@pytest.mark.gpu
@pytest.mark.parametrize("sa_dtype", [np.uint32, np.uint64])
def test_device_index_fm_header_ranges_and_rows(hip, sa_dtype):
    lib = index_lib()
    text = synth.repeat_text(5, 300000, contigs=4)  # past 2^16 suffixes: a prefix directory
    n = text.size
    sa = oracle.fast_build(text)[0].astype(sa_dtype)
    d_text, d_sa = hip.DeviceArray(n + 64), hip.DeviceArray(sa.nbytes)
    d_text.upload(text)
    d_sa.upload(sa)
    index = ctypes.c_void_p()
    err = ctypes.create_string_buffer(256)
    assert lib.sfk_hip_index_create(0, d_text.ptr, n, d_sa.ptr, 8 * sa.itemsize, ctypes.byref(index), err, 256) == 0
    try:
        lib.sfk_hip_index_prefix_k.argtypes = [ctypes.c_void_p]
        lib.sfk_hip_index_prefix_k.restype = ctypes.c_uint32
        assert lib.sfk_hip_index_prefix_k(index) > 0
        assert lib.sfk_hip_index_build_fm(index, err, 256) == 0, err.value
        header = (ctypes.c_uint64 * 6)()
        assert lib.sfk_hip_index_fm_header(index, header) == 0
        counts = np.bincount(text, minlength=7)
        low = int(counts[0] + counts[1])
        expected = [low]
        for c in range(2, 6):
            expected.append(expected[-1] + int(counts[c]))
        assert list(header)[:5] == expected and header[5] == low
        # Ranges of mixed-orientation patterns against the oracle.
        rng = np.random.default_rng(4)
        plain = text[(text >= 2) & (text <= 5)]
        pats, orient = [], []
        for length in (1, 4, 9, 10, 11, 12, 30, 45):
            for o in (0, 1):
                at = int(rng.integers(0, plain.size - length))
                pats.append((plain[at:at + length] - 2).astype(np.uint8))
                orient.append(o)
        codes = np.concatenate(pats)
        offsets = np.concatenate([[0], np.cumsum([p.size for p in pats])]).astype(np.uint64)
        d_codes, d_off = hip.DeviceArray(codes.nbytes), hip.DeviceArray(offsets.nbytes)
        d_or, d_rng = hip.DeviceArray(len(pats)), hip.DeviceArray(16 * len(pats))
        d_codes.upload(codes)
        d_off.upload(offsets)
        d_or.upload(np.array(orient, np.uint8))
        assert lib.sfk_hip_index_fm_ranges(index, d_codes.ptr, d_off.ptr, d_or.ptr, len(pats), d_rng.ptr, None) == 0
        ranges = d_rng.download(np.uint64, 2 * len(pats)).reshape(-1, 2)
        for p, o, r in zip(pats, orient, ranges):
            s = "".join("ACGT"[c] for c in p)
            if o:
                s = oracle.reverse_complement(s)
            assert tuple(int(x) for x in r) == oracle.equal_range(text, s), (s, o)
            if r[1] > r[0]:
                rows = (ctypes.c_uint64 * int(r[1] - r[0]))()
                assert lib.sfk_hip_index_sa_rows(index, int(r[0]), int(r[1]), rows) == 0
                assert list(rows) == [int(v) for v in sa[int(r[0]):int(r[1])]]
        # FM counts equal the SA-search counts on the served query mix.
        q = 20000
        d_p, d_l = hip.DeviceArray(8 * q), hip.DeviceArray(q)
        d_c1, d_c2 = hip.DeviceArray(4 * q), hip.DeviceArray(4 * q)
        assert lib.sfk_hip_index_generate_queries(index, 99, q, 1, 32, d_p.ptr, d_l.ptr, None) == 0
        assert lib.sfk_hip_index_count(index, d_p.ptr, d_l.ptr, q, d_c1.ptr, None) == 0
        assert lib.sfk_hip_index_fm_count(index, d_p.ptr, d_l.ptr, q, d_c2.ptr, None) == 0
        c1, c2 = d_c1.download(np.uint32, q), d_c2.download(np.uint32, q)
        assert np.array_equal(c1, c2) and c1.sum() > 0
    finally:
        lib.sfk_hip_index_destroy(index)


@pytest.mark.gpu
def test_device_fm_save_is_unsupported(hip, tmp_path):
    from sufkit_amd import SufkitError

    fm = device_fm([("c", b"ACGTACGT")])
    with pytest.raises(SufkitError) as e:
        fm.save(tmp_path / "x.sufidx")
    assert e.value.code == "unsupported_backend"


@pytest.mark.gpu
def test_host_fm_index_with_a_prefix_directory(hip):
    rng = np.random.default_rng(21)
    seqs = [synth_bases(rng, 130000, n_runs=3), synth_bases(rng, 60000, n_runs=1)]  # k = 8 from 161K symbols
    fm = device_fm([("a", seqs[0]), ("b", seqs[1])])
    summary = fm.summary()
    assert summary["prefix_directory_k"] > 0 and summary["sa_bits"] == 32
    text, _ = oracle.encode_records(seqs)
    starts = np.array([0, len(seqs[0]) + 1], np.uint64)
    k = summary["prefix_directory_k"]
    plain = text[(text >= 2) & (text <= 5)]
    for length in (k - 1, k, k + 1, 2 * k + 3):
        for _ in range(5):
            at = int(rng.integers(0, plain.size - length))
            p = "".join("ACGT"[c - 2] for c in plain[at:at + length])
            assert fm.equal_range(p) == oracle.equal_range(text, p), p
            assert as_hits(fm.locate(p, 2)) == as_hits(oracle.locate(text, starts, p, 2)), p


@pytest.mark.gpu
@pytest.mark.parametrize("request_index", [6, 15, 1])
def test_fm_counts_equal_sa_counts_on_a_served_request(hip, request_index):
    # The three smallest requests of the served workload (1.0-1.6 M symbols,
    # prefix directory k = 9), each with its million generated count queries.
    lib = index_lib()
    text, _ = hip.synthetic_text(5, 20260822, request_index, 1.0)
    n = text.size
    sa, _, _, _ = hip.hip_build(text)
    d_text, d_sa = hip.DeviceArray(n + 64), hip.DeviceArray(sa.nbytes)
    d_text.upload(text)
    d_sa.upload(sa)
    index = ctypes.c_void_p()
    err = ctypes.create_string_buffer(256)
    assert lib.sfk_hip_index_create(0, d_text.ptr, n, d_sa.ptr, 32, ctypes.byref(index), err, 256) == 0
    try:
        assert lib.sfk_hip_index_build_fm(index, err, 256) == 0, err.value
        q = 1_000_000
        d_p, d_l = hip.DeviceArray(8 * q), hip.DeviceArray(q)
        d_c1, d_c2 = hip.DeviceArray(4 * q), hip.DeviceArray(4 * q)
        seed = 7 ^ ((request_index * 0x9E3779B97F4A7C15) & ((1 << 64) - 1))
        assert lib.sfk_hip_index_generate_queries(index, seed, q, 20, 32, d_p.ptr, d_l.ptr, None) == 0
        assert lib.sfk_hip_index_count(index, d_p.ptr, d_l.ptr, q, d_c1.ptr, None) == 0
        assert lib.sfk_hip_index_fm_count(index, d_p.ptr, d_l.ptr, q, d_c2.ptr, None) == 0
        c1, c2 = d_c1.download(np.uint32, q), d_c2.download(np.uint32, q)
        bad = np.nonzero(c1 != c2)[0]
        if bad.size:
            pats, lens = d_p.download(np.uint64, q), d_l.download(np.uint8, q)
            s = "".join("$|ACGTN"[c] for c in text.tolist())
            for b in bad[:5].tolist():
                p = hip.decode_pattern(int(pats[b]), int(lens[b]))
                want, at = 0, s.find(p)
                while at != -1:
                    want += 1
                    at = s.find(p, at + 1)
                print(f"query {b}: {p} (len {lens[b]}) sa {c1[b]} fm {c2[b]} scan {want}", flush=True)
        assert bad.size == 0, f"{bad.size} FM counts differ from SA counts"
    finally:
        lib.sfk_hip_index_destroy(index)

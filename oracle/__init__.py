"""CPU oracle for the suffix-index construction path.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, and only as the checker; the product
(sufkit_amd/) never imports it.

Two layers:
  * liboracle.so (oracle.c): a plain-C restatement of the reference's
    normalization, SA order, ISA scatter, generalised Kasai LCP, FNV-1a-64
    and CRC32, each citing the reference lines it follows;
  * _ref/libsufkit_ref.so: the reference's own divsufsort backend
    (detail::BuildDivsufsort32/64 over libdivsufsort 2.0.2) compiled from
    the reference sources by oracle/Makefile, when those sources were
    present at build time;
  * _ref/libsufkit_refcaps.so: the reference's CaPS-SA backend
    (detail::BuildCaps32/64, multi-threaded) compiled the same way, the
    multi-core CPU baseline.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_LIB = None
_REF = None

U64P = ctypes.POINTER(ctypes.c_uint64)
U8P = ctypes.POINTER(ctypes.c_uint8)


def build(ref: str | None = "/root/reference") -> None:
    """Compile liboracle.so and, when the reference sources exist, _ref/."""
    env = dict(os.environ)
    args = ["make", "-s", "-C", str(HERE), "oracle"]
    subprocess.run(args, check=True, env=env)
    if ref and Path(ref, "src", "divsufsort_backend.cpp").exists():
        subprocess.run(["make", "-s", "-j8", "-C", str(HERE), "ref", "ref-caps", "ref-full", f"REF={ref}"],
                       check=True, env=env)


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        path = HERE / "liboracle.so"
        if not path.exists():
            build(ref=None)
        L = ctypes.CDLL(str(path))
        L.sfk_oracle_encode_records.argtypes = [U8P, U64P, ctypes.c_uint64, U8P, U64P]
        L.sfk_oracle_encode_records.restype = ctypes.c_uint64
        L.sfk_oracle_fingerprint.argtypes = [U8P, ctypes.c_uint64]
        L.sfk_oracle_fingerprint.restype = ctypes.c_uint64
        L.sfk_oracle_sa_naive.argtypes = [U8P, ctypes.c_uint64, U64P]
        L.sfk_oracle_sa_naive.restype = None
        L.sfk_oracle_sa_doubling.argtypes = [U8P, ctypes.c_uint64, U64P]
        L.sfk_oracle_sa_doubling.restype = ctypes.c_int
        L.sfk_oracle_isa.argtypes = [U64P, ctypes.c_uint64, ctypes.c_uint32, U64P]
        L.sfk_oracle_isa.restype = None
        L.sfk_oracle_lcp_kasai.argtypes = [U8P, ctypes.c_uint64, U64P, U64P, ctypes.c_uint64,
                                           ctypes.c_uint32, U64P]
        L.sfk_oracle_lcp_kasai.restype = None
        L.sfk_oracle_lcp_brute.argtypes = [U8P, ctypes.c_uint64, U64P, ctypes.c_uint64, U64P]
        L.sfk_oracle_lcp_brute.restype = None
        L.sfk_oracle_crc32.argtypes = [ctypes.c_uint32, U8P, ctypes.c_uint64]
        L.sfk_oracle_crc32.restype = ctypes.c_uint32
        L.sfk_oracle_digest_u64.argtypes = [U64P, ctypes.c_uint64]
        L.sfk_oracle_digest_u64.restype = ctypes.c_uint64
        L.sfk_oracle_sum_digest_u64.argtypes = [U64P, ctypes.c_uint64]
        L.sfk_oracle_sum_digest_u64.restype = ctypes.c_uint64
        L.sfk_oracle_sum_digest_u32.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint64]
        L.sfk_oracle_sum_digest_u32.restype = ctypes.c_uint64
        L.sfk_oracle_lean_digests.argtypes = [U8P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint32), U64P]
        L.sfk_oracle_lean_digests.restype = ctypes.c_int
        L.sfk_oracle_prefix_directory.argtypes = [U8P, ctypes.c_uint64, U64P, ctypes.c_uint32, U64P, U64P, U64P]
        L.sfk_oracle_prefix_directory.restype = ctypes.c_uint64
        _LIB = L
    return _LIB


def ref_available() -> bool:
    return (HERE / "_ref" / "libsufkit_ref.so").exists()


def ref() -> ctypes.CDLL:
    global _REF
    if _REF is None:
        path = HERE / "_ref" / "libsufkit_ref.so"
        if not path.exists():
            raise FileNotFoundError(f"{path} not built (reference sources absent at build time)")
        R = ctypes.CDLL(str(path))
        R.sfk_ref_build_fast.argtypes = [U8P, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint32,
                                         ctypes.c_uint32, U64P, U64P, U64P,
                                         ctypes.POINTER(ctypes.c_double)]
        R.sfk_ref_build_fast.restype = ctypes.c_longlong
        R.sfk_ref_divsufsort.argtypes = [U8P, ctypes.c_uint64, ctypes.c_int, U64P]
        R.sfk_ref_divsufsort.restype = ctypes.c_int
        _REF = R
    return _REF


def prefix_select_k(suffix_count: int, resident_core_bytes: int) -> int:
    """FastPrefixIndex::SelectK with default options, restated
    (fast_prefix_index.cpp:286-307): none below 2^16 suffixes, else the
    largest k in [8, 10] whose 4^k entries of two row-width coordinates fit
    25% of the resident core bytes."""
    if suffix_count < (1 << 16):
        return 0
    budget = (resident_core_bytes // 10000) * 2500 + ((resident_core_bytes % 10000) * 2500) // 10000
    row_bytes = 4 if suffix_count <= 0xFFFFFFFF else 8
    for k in (10, 9, 8):
        if (1 << (2 * k)) * 2 * row_bytes <= budget:
            return k
    return 0


def prefix_directory(text: np.ndarray, isa_: np.ndarray, k: int):
    """(begin, end, indexed k-mers, non-empty entries) of the Fast prefix
    directory, from the restatement in oracle.c."""
    text, p = _u8(text)
    isa_ = np.ascontiguousarray(isa_, dtype=np.uint64)
    begin = np.empty(1 << (2 * k), np.uint64)
    end = np.empty(1 << (2 * k), np.uint64)
    nonempty = np.zeros(1, np.uint64)
    indexed = lib().sfk_oracle_prefix_directory(p, text.size, _u64(isa_), k, _u64(begin), _u64(end), _u64(nonempty))
    return begin, end, int(indexed), int(nonempty[0])


def contigs_of(text: np.ndarray):
    """Contig starts and lengths of a logical text (separators end contigs;
    the last contig ends before the sentinel)."""
    seps = np.flatnonzero(np.asarray(text) == 1)
    ends = np.concatenate([seps, [text.size - 1]]) if (seps.size == 0 or seps[-1] != text.size - 2) else seps
    starts = np.concatenate([[0], ends[:-1] + 1]).astype(np.uint64)
    return starts, (ends - starts).astype(np.uint64)


def ref_prefix_directory(text: np.ndarray, sa: np.ndarray, isa_: np.ndarray, resident_core_bytes: int):
    """The reference's own FastPrefixIndex over the same arrays: (k, begin,
    end, indexed, nonempty), or None when the reference writer library was
    not built."""
    if not ref_writer_available():
        return None
    W = ctypes.CDLL(str(HERE / "_ref" / "libsufkit_refwriter.so"))
    W.sfk_ref_prefix_directory.argtypes = [U8P, ctypes.c_uint64, U64P, U64P, U64P, U64P, ctypes.c_uint64,
                                           ctypes.c_uint64, U64P, U64P, U64P]
    W.sfk_ref_prefix_directory.restype = ctypes.c_int
    text, p = _u8(text)
    sa = np.ascontiguousarray(sa, np.uint64)
    isa_ = np.ascontiguousarray(isa_, np.uint64)
    starts, lengths = contigs_of(text)
    begin = np.empty(1 << 20, np.uint64)
    end = np.empty(1 << 20, np.uint64)
    counts = np.zeros(2, np.uint64)
    k = W.sfk_ref_prefix_directory(p, text.size, _u64(sa), _u64(isa_), _u64(starts), _u64(lengths), starts.size,
                                   resident_core_bytes, _u64(begin), _u64(end), _u64(counts))
    if k < 0:
        raise RuntimeError("reference FastPrefixIndex::Build failed")
    return k, begin[: 1 << (2 * k)], end[: 1 << (2 * k)], int(counts[0]), int(counts[1])


def ref_prefix_select_k(suffix_count: int, resident_core_bytes: int):
    if not ref_writer_available():
        return None
    W = ctypes.CDLL(str(HERE / "_ref" / "libsufkit_refwriter.so"))
    W.sfk_ref_prefix_select_k.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    W.sfk_ref_prefix_select_k.restype = ctypes.c_uint32
    return int(W.sfk_ref_prefix_select_k(suffix_count, resident_core_bytes))


_CAPS = None


def ref_caps_available() -> bool:
    return (HERE / "_ref" / "libsufkit_refcaps.so").exists()


def ref_caps_fast(text: np.ndarray, width_bits: int, threads: int):
    """SA, ISA, LCP (uint64) from the reference's CaPS-SA backend with
    `threads` threads, and its (SA+LCP, ISA) seconds."""
    global _CAPS
    if _CAPS is None:
        path = HERE / "_ref" / "libsufkit_refcaps.so"
        if not path.exists():
            raise FileNotFoundError(f"{path} not built (reference sources absent at build time)")
        C = ctypes.CDLL(str(path))
        C.sfk_ref_caps_fast.argtypes = [U8P, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint32, U64P, U64P, U64P,
                                        ctypes.POINTER(ctypes.c_double)]
        C.sfk_ref_caps_fast.restype = ctypes.c_longlong
        _CAPS = C
    text, p = _u8(text)
    n = text.size
    sa, isa_, lcp = np.empty(n, np.uint64), np.empty(n, np.uint64), np.empty(n, np.uint64)
    secs = (ctypes.c_double * 2)()
    rows = _CAPS.sfk_ref_caps_fast(p, n, width_bits, threads, _u64(sa), _u64(isa_), _u64(lcp), secs)
    if rows != n:
        raise RuntimeError("reference CaPS-SA build failed")
    return sa, isa_, lcp, (secs[0], secs[1])


def _u8(a):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    return a, a.ctypes.data_as(U8P)


def _u64(a):
    return a.ctypes.data_as(U64P)


def encode_records(sequences: list[bytes]) -> tuple[np.ndarray, int]:
    """Logical text of FromRecords + sentinel; returns (text, ambiguous)."""
    raw = np.frombuffer(b"".join(sequences), dtype=np.uint8)
    raw, raw_p = _u8(raw if raw.size else np.zeros(1, np.uint8))
    lengths = np.array([len(s) for s in sequences], dtype=np.uint64)
    out = np.empty(int(lengths.sum()) + len(sequences) + 1, dtype=np.uint8)
    amb = ctypes.c_uint64(0)
    n = lib().sfk_oracle_encode_records(raw_p, _u64(lengths), len(sequences),
                                        out.ctypes.data_as(U8P), ctypes.byref(amb))
    assert n == out.size
    return out, int(amb.value)


def fingerprint(data: np.ndarray) -> int:
    data, p = _u8(data)
    return int(lib().sfk_oracle_fingerprint(p, data.size))


def sa_naive(text: np.ndarray) -> np.ndarray:
    text, p = _u8(text)
    sa = np.empty(text.size, dtype=np.uint64)
    lib().sfk_oracle_sa_naive(p, text.size, _u64(sa))
    return sa


def sa_doubling(text: np.ndarray) -> np.ndarray:
    text, p = _u8(text)
    sa = np.empty(text.size, dtype=np.uint64)
    if lib().sfk_oracle_sa_doubling(p, text.size, _u64(sa)) != 0:
        raise MemoryError("oracle doubling allocation failed")
    return sa


def isa(sa: np.ndarray, sampling_rate: int = 1) -> np.ndarray:
    sa = np.ascontiguousarray(sa, dtype=np.uint64)
    out = np.empty(sa.size, dtype=np.uint64)
    lib().sfk_oracle_isa(_u64(sa), sa.size, sampling_rate, _u64(out))
    return out


def lcp_kasai(text: np.ndarray, sa: np.ndarray, isa_: np.ndarray, sampling_rate: int = 1) -> np.ndarray:
    text, p = _u8(text)
    sa = np.ascontiguousarray(sa, dtype=np.uint64)
    isa_ = np.ascontiguousarray(isa_, dtype=np.uint64)
    out = np.empty(sa.size, dtype=np.uint64)
    lib().sfk_oracle_lcp_kasai(p, text.size, _u64(sa), _u64(isa_), sa.size, sampling_rate, _u64(out))
    return out


def lcp_brute(text: np.ndarray, sa: np.ndarray) -> np.ndarray:
    text, p = _u8(text)
    sa = np.ascontiguousarray(sa, dtype=np.uint64)
    out = np.empty(sa.size, dtype=np.uint64)
    lib().sfk_oracle_lcp_brute(p, text.size, _u64(sa), sa.size, _u64(out))
    return out


def crc32(data: bytes | np.ndarray, crc: int = 0) -> int:
    arr = np.frombuffer(data, dtype=np.uint8) if isinstance(data, (bytes, bytearray)) else data
    arr, p = _u8(arr if arr.size else np.zeros(1, np.uint8))
    size = len(data) if isinstance(data, (bytes, bytearray)) else data.size
    return int(lib().sfk_oracle_crc32(crc, p, size))


def digest(values: np.ndarray) -> int:
    values = np.ascontiguousarray(values, dtype=np.uint64)
    return int(lib().sfk_oracle_digest_u64(_u64(values), values.size))


def sum_digest(values: np.ndarray) -> int:
    """Position-keyed sum digest (oracle.c sfk_oracle_sum_digest_u64): the
    form a device reduction can compute."""
    if values.dtype == np.uint32:
        values = np.ascontiguousarray(values)
        return int(lib().sfk_oracle_sum_digest_u32(
            values.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), values.size))
    values = np.ascontiguousarray(values, dtype=np.uint64)
    return int(lib().sfk_oracle_sum_digest_u64(_u64(values), values.size))


def lean_digests(text: np.ndarray, sa32: np.ndarray) -> dict:
    """Sum digests of SA, ISA and LCP from a complete 32-bit SA (ISA scatter
    and Phi/PLCP on the CPU, 8n bytes of work arrays)."""
    text, p = _u8(text)
    sa32 = np.ascontiguousarray(sa32, dtype=np.uint32)
    out = np.zeros(4, np.uint64)
    rc = lib().sfk_oracle_lean_digests(p, text.size, sa32.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                       _u64(out))
    if rc != 0:
        raise RuntimeError("lean digests: SA is not a permutation or memory ran out")
    return {"sa": int(out[0]), "isa": int(out[1]), "lcp": int(out[2]), "lcp_max": int(out[3])}


def ref_divsufsort64_u32(text: np.ndarray) -> tuple[np.ndarray, float]:
    """The compiled reference's divsufsort64 SA, narrowed to 32 bits (texts
    of up to 2^32 - 1 symbols); returns (SA, seconds of the divsufsort64 call)."""
    text, p = _u8(text)
    sa = np.empty(text.size, np.uint32)
    secs = (ctypes.c_double * 1)()
    R = ref()
    if not hasattr(R, "_narrow_set"):
        R.sfk_ref_divsufsort64_narrow.argtypes = [U8P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint32),
                                                  ctypes.POINTER(ctypes.c_double)]
        R.sfk_ref_divsufsort64_narrow.restype = ctypes.c_int
        R._narrow_set = True
    if R.sfk_ref_divsufsort64_narrow(p, text.size, sa.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), secs) != 0:
        raise RuntimeError("reference divsufsort64 failed")
    return sa, float(secs[0])


def fast_build(text: np.ndarray, method: str = "auto"):
    """SA, ISA, LCP (uint64) of the Fast profile.

    method "ref32"/"ref64": the compiled reference (BuildDivsufsort32/64);
    "doubling"/"naive": the restatement; "auto": reference if built, else
    doubling.
    """
    text, p = _u8(text)
    n = text.size
    if method == "auto":
        method = ("ref32" if n <= 0x7FFFFFFF else "ref64") if ref_available() else "doubling"
    if method in ("ref32", "ref64"):
        sa = np.empty(n, np.uint64)
        isa_ = np.empty(n, np.uint64)
        lcp = np.empty(n, np.uint64)
        secs = (ctypes.c_double * 3)()
        rows = ref().sfk_ref_build_fast(p, n, 32 if method == "ref32" else 64, 1, 1,
                                        _u64(sa), _u64(isa_), _u64(lcp), secs)
        if rows != n:
            raise RuntimeError("reference build failed")
        return sa, isa_, lcp
    sa = sa_naive(text) if method == "naive" else sa_doubling(text)
    isa_ = isa(sa)
    return sa, isa_, lcp_kasai(text, sa, isa_)


def ref_build_sampled(text: np.ndarray, sampling_rate: int, width_bits: int = 32):
    """The compiled reference's BuildDivsufsort32/64 with sampling rate K:
    (sampled SA, ISA over positions / K, LCP) as uint64."""
    text, p = _u8(text)
    n = text.size
    rows = (n + sampling_rate - 1) // sampling_rate
    sa = np.empty(n, np.uint64)
    isa_ = np.empty(n, np.uint64)
    lcp = np.empty(n, np.uint64)
    secs = (ctypes.c_double * 3)()
    got = ref().sfk_ref_build_fast(p, n, width_bits, sampling_rate, 1, _u64(sa), _u64(isa_), _u64(lcp), secs)
    if got != rows:
        raise RuntimeError("reference sampled build failed")
    return sa[:rows].copy(), isa_[:rows].copy(), lcp[:rows].copy()


def ref_writer_available() -> bool:
    return (HERE / "_ref" / "libsufkit_refwriter.so").exists()


def ref_write_sufidx(records: list[tuple[str, bytes]], path: str, coordinate_width: int = 0,
                     storage_width: int = 0, acceleration: int = 3, resource_profile: int = 0) -> None:
    """Build with the reference's own SuffixArray::Build (divsufsort) and Save
    a .sufidx 1.4 file (test infrastructure: golden files only)."""
    W = ctypes.CDLL(str(HERE / "_ref" / "libsufkit_refwriter.so"))
    W.sfk_ref_write_sufidx_profile.argtypes = [ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_char_p),
                                               U64P, ctypes.c_uint64, ctypes.c_char_p, ctypes.c_int,
                                               ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               ctypes.POINTER(ctypes.c_double), ctypes.c_char_p, ctypes.c_size_t]
    W.sfk_ref_write_sufidx_profile.restype = ctypes.c_int
    names = (ctypes.c_char_p * len(records))(*[r[0].encode() for r in records])
    seqs = (ctypes.c_char_p * len(records))(*[r[1] for r in records])
    lens = np.array([len(r[1]) for r in records], dtype=np.uint64)
    err = ctypes.create_string_buffer(512)
    rc = W.sfk_ref_write_sufidx_profile(names, seqs, _u64(lens), len(records), str(path).encode(),
                                        coordinate_width, storage_width, acceleration, resource_profile, None,
                                        err, len(err))
    if rc != 0:
        raise RuntimeError(err.value.decode())


def ref_write_sufidx_timed(records: list[tuple[str, bytes]], path: str, coordinate_width: int = 0,
                           storage_width: int = 0, acceleration: int = 3) -> dict:
    """ref_write_sufidx, returning the wall time of the reference's own
    normalization, Build and Save (the CPU baseline of saving an index)."""
    W = ctypes.CDLL(str(HERE / "_ref" / "libsufkit_refwriter.so"))
    W.sfk_ref_write_sufidx_timed.argtypes = [ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_char_p),
                                             U64P, ctypes.c_uint64, ctypes.c_char_p, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                             ctypes.c_char_p, ctypes.c_size_t]
    W.sfk_ref_write_sufidx_timed.restype = ctypes.c_int
    names = (ctypes.c_char_p * len(records))(*[r[0].encode() for r in records])
    seqs = (ctypes.c_char_p * len(records))(*[r[1] for r in records])
    lens = np.array([len(r[1]) for r in records], dtype=np.uint64)
    seconds = (ctypes.c_double * 3)()
    err = ctypes.create_string_buffer(512)
    rc = W.sfk_ref_write_sufidx_timed(names, seqs, _u64(lens), len(records), str(path).encode(),
                                      coordinate_width, storage_width, acceleration, seconds, err, len(err))
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return dict(normalize_seconds=seconds[0], build_seconds=seconds[1], save_seconds=seconds[2])


# ------------------------------------------------------------ exact search

_BASE_CODE = {ord(c): v for c, v in zip("ACGTacgt", (2, 3, 4, 5, 2, 3, 4, 5))}


def encode_pattern(pattern: str) -> np.ndarray:
    """A/C/G/T pattern as text codes 2..5 (query.cpp:47-65)."""
    return np.array([_BASE_CODE[ord(ch)] for ch in pattern], dtype=np.uint8)


def reverse_complement(pattern: str) -> str:
    return pattern[::-1].translate(str.maketrans("ACGTacgt", "TGCAtgca"))


def occurrences(text: np.ndarray, pattern: str) -> np.ndarray:
    """Positions where the pattern occurs in the text, increasing."""
    L = lib()
    L.sfk_oracle_occurrences.argtypes = [U8P, ctypes.c_uint64, U8P, ctypes.c_uint64, U64P, ctypes.c_uint64]
    L.sfk_oracle_occurrences.restype = ctypes.c_uint64
    text, tp = _u8(text)
    pat, pp = _u8(encode_pattern(pattern))
    total = int(L.sfk_oracle_occurrences(tp, text.size, pp, pat.size, None, 0))
    out = np.zeros(max(total, 1), np.uint64)
    L.sfk_oracle_occurrences(tp, text.size, pp, pat.size, _u64(out), total)
    return out[:total]


def equal_range(text: np.ndarray, pattern: str) -> tuple[int, int]:
    """The pattern's SA row range [begin, end) by definition, (0, 0) when absent."""
    L = lib()
    L.sfk_oracle_suffixes_below.argtypes = [U8P, ctypes.c_uint64, U8P, ctypes.c_uint64]
    L.sfk_oracle_suffixes_below.restype = ctypes.c_uint64
    count = occurrences(text, pattern).size
    if count == 0:
        return 0, 0
    text, tp = _u8(text)
    pat, pp = _u8(encode_pattern(pattern))
    below = int(L.sfk_oracle_suffixes_below(tp, text.size, pp, pat.size))
    return below, below + count


def locate(text: np.ndarray, contig_starts: np.ndarray, pattern: str, strands: int = 0,
           max_hits: int | None = None) -> dict:
    """FmIndex::Locate by definition (fm_index.cpp:112-181,612-677):
    (sequence id, local position, length, strand) sorted by global position,
    a position found on both strands of a palindrome reported once as both
    (strand 2), the max_hits smallest kept."""
    rc = reverse_complement(pattern)
    if strands == 0:
        lanes = [(pattern, 0)]
    elif strands == 1:
        lanes = [(rc, 1)]
    else:
        lanes = [(pattern, 2)] if rc.upper() == pattern.upper() else [(pattern, 0), (rc, 1)]
    hits = sorted((int(p), s) for lane, s in lanes for p in occurrences(text, lane))
    total = len(hits)
    if max_hits is not None:
        hits = hits[:max_hits]
    starts = np.asarray(contig_starts, np.uint64)
    out = []
    for pos, s in hits:
        cid = int(np.searchsorted(starts, pos, side="right")) - 1
        out.append((cid, pos - int(starts[cid]), len(pattern), s))
    return {"total_hits": total, "hits": out, "truncated": len(out) < total}


def ref_fm_available() -> bool:
    return ref_writer_available()


class RefFmIndex:
    """The reference's own FmIndex (SDSL csa_wt on the host), compiled into
    oracle/_ref; test infrastructure only (golden vectors, error codes)."""

    def __init__(self, records: list[tuple[str, bytes]], backend: int = 1):
        W = ctypes.CDLL(str(HERE / "_ref" / "libsufkit_refwriter.so"))
        vp, cp = ctypes.c_void_p, ctypes.c_char_p
        W.sfk_ref_fm_build.argtypes = [ctypes.POINTER(cp), ctypes.POINTER(cp), U64P, ctypes.c_uint64, ctypes.c_int,
                                       ctypes.POINTER(vp)]
        W.sfk_ref_fm_free.argtypes = [vp]
        W.sfk_ref_fm_free.restype = None
        W.sfk_ref_fm_equal_range.argtypes = [vp, cp, ctypes.c_uint64, U64P]
        W.sfk_ref_fm_count.argtypes = [vp, cp, ctypes.c_uint64, ctypes.c_int, U64P]
        W.sfk_ref_fm_count_batch.argtypes = [vp, cp, U64P, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint32, U64P]
        W.sfk_ref_fm_locate.argtypes = [vp, cp, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, U64P,
                                        ctypes.POINTER(ctypes.c_uint32), U64P, U64P, U8P, ctypes.c_uint64]
        self._W = W
        names = (cp * len(records))(*[r[0].encode() for r in records])
        seqs = (cp * len(records))(*[bytes(r[1]) for r in records])
        lens = np.array([len(r[1]) for r in records], dtype=np.uint64)
        self._h = vp()
        self.build_status = W.sfk_ref_fm_build(names, seqs, _u64(lens), len(records), backend, ctypes.byref(self._h))

    def equal_range(self, pattern: str):
        out = (ctypes.c_uint64 * 2)()
        p = pattern.encode()
        rc = self._W.sfk_ref_fm_equal_range(self._h, p, len(p), out)
        return (int(out[0]), int(out[1])) if rc == 0 else ("error", rc)

    def count(self, pattern: str, strands: int = 0):
        out = ctypes.c_uint64()
        p = pattern.encode()
        rc = self._W.sfk_ref_fm_count(self._h, p, len(p), strands, ctypes.byref(out))
        return int(out.value) if rc == 0 else ("error", rc)

    def count_batch(self, patterns: list[str], strands: int = 0, batch_width: int = 0):
        data = [p.encode() for p in patterns]
        offsets = np.zeros(len(data) + 1, np.uint64)
        if data:
            offsets[1:] = np.cumsum([len(d) for d in data])
        out = np.zeros(max(len(data), 1), np.uint64)
        rc = self._W.sfk_ref_fm_count_batch(self._h, b"".join(data) or b"\0", _u64(offsets), len(data), strands,
                                            batch_width, _u64(out))
        return [int(v) for v in out[:len(data)]] if rc == 0 else ("error", rc)

    def locate(self, pattern: str, strands: int = 0, max_hits: int | None = None):
        info = (ctypes.c_uint64 * 3)()
        p = pattern.encode()
        cap = 1 << 16
        ids = np.zeros(cap, np.uint32)
        pos = np.zeros(cap, np.uint64)
        lens = np.zeros(cap, np.uint64)
        strs = np.zeros(cap, np.uint8)
        rc = self._W.sfk_ref_fm_locate(self._h, p, len(p), strands, int(max_hits is not None), max_hits or 0, info,
                                       ids.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), _u64(pos), _u64(lens),
                                       strs.ctypes.data_as(U8P), cap)
        if rc != 0:
            return ("error", rc)
        size = min(int(info[1]), cap)
        hits = [(int(ids[i]), int(pos[i]), int(lens[i]), int(strs[i])) for i in range(size)]
        return {"total_hits": int(info[0]), "hits": hits, "truncated": bool(info[2])}

    def __del__(self):
        if getattr(self, "_h", None):
            self._W.sfk_ref_fm_free(self._h)
            self._h = None

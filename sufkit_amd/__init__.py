"""sufkit_amd: MI355X-native suffix-index construction (SA + ISA + raw LCP).

The product is native code: the HIP/gfx950 backend behind the C ABI in
include/sufkit_hip.h (sufkit_amd/lib/libsufkit_hip.so) and the C++17 host
library around it (sufkit_amd/lib/libsufkit.so).  This module only loads
them with ctypes and converts numpy buffers; it has no fallback: when the
native library is missing every call raises.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_DIR = PKG / "lib"

SFK_HIP_RETAIN_ISA = 1
SFK_HIP_RETAIN_LCP = 2
SFK_HIP_PROFILE_KERNELS = 4

# Reference ErrorCode + 1 (include/sufkit/types.hpp:708-722).
STATUS_NAMES = {
    0: "ok",
    1: "invalid_input",
    2: "io_error",
    3: "unsupported_backend",
    4: "corrupt_index",
    5: "version_mismatch",
    6: "build_failure",
}


class SufkitError(RuntimeError):
    """Raised for a non-zero status; `code` is the reference ErrorCode name."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.code = STATUS_NAMES.get(status, str(status))


class HipStats(ctypes.Structure):
    _fields_ = [
        ("sa_seconds", ctypes.c_double),
        ("isa_seconds", ctypes.c_double),
        ("lcp_seconds", ctypes.c_double),
        ("device_seconds", ctypes.c_double),
        ("h2d_seconds", ctypes.c_double),
        ("d2h_seconds", ctypes.c_double),
        ("refine_steps", ctypes.c_uint64),
        ("doubling_rounds", ctypes.c_uint64),
        ("active_total", ctypes.c_uint64),
        ("sort_passes", ctypes.c_uint64),
        ("lce_long_items", ctypes.c_uint64),
        ("lce_long_words", ctypes.c_uint64),
        ("algorithmic_bytes_sa", ctypes.c_uint64),
        ("algorithmic_bytes_isa", ctypes.c_uint64),
        ("algorithmic_bytes_lcp", ctypes.c_uint64),
        ("sort_scatter_seconds", ctypes.c_double),
        ("sort_scatter_launches", ctypes.c_uint64),
        ("sort_scatter_bytes", ctypes.c_uint64),
        ("prefix_directory_seconds", ctypes.c_double),
        ("call_seconds", ctypes.c_double),
        ("lcp_gathered_rows", ctypes.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_HIP = None


def hip_library_path() -> Path:
    return LIB_DIR / "libsufkit_hip.so"


def load_hip() -> ctypes.CDLL:
    """Load libsufkit_hip.so (built by sufkit_amd.build); raise if absent."""
    global _HIP
    if _HIP is not None:
        return _HIP
    path = hip_library_path()
    if not path.exists():
        raise RuntimeError(
            f"{path} is missing: build it with `python -m sufkit_amd.build` "
            "(there is no CPU fallback)")
    lib = ctypes.CDLL(str(path))
    u8p = ctypes.POINTER(ctypes.c_uint8)
    vp = ctypes.c_void_p
    lib.sfk_hip_abi_version.restype = ctypes.c_int
    lib.sfk_hip_device_count.restype = ctypes.c_int
    lib.sfk_hip_builder_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp), ctypes.c_char_p, ctypes.c_size_t]
    lib.sfk_hip_builder_create.restype = ctypes.c_int
    lib.sfk_hip_builder_destroy.argtypes = [vp]
    lib.sfk_hip_builder_destroy.restype = None
    lib.sfk_hip_builder_stream.argtypes = [vp]
    lib.sfk_hip_builder_stream.restype = vp
    lib.sfk_hip_workspace_bytes.argtypes = [ctypes.c_uint64, ctypes.c_uint32]
    lib.sfk_hip_workspace_bytes.restype = ctypes.c_uint64
    lib.sfk_hip_build_device.argtypes = [
        vp, vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
        vp, ctypes.c_uint32, vp, ctypes.c_uint32, vp, ctypes.c_uint32,
        ctypes.POINTER(HipStats), ctypes.c_char_p, ctypes.c_size_t]
    lib.sfk_hip_build_device.restype = ctypes.c_int
    lib.sfk_hip_build_batch_device.argtypes = [
        vp, vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint32,
        ctypes.c_uint32, ctypes.c_uint32, vp, ctypes.c_uint32, vp, ctypes.c_uint32, vp,
        ctypes.c_uint32, ctypes.POINTER(HipStats), ctypes.c_char_p, ctypes.c_size_t]
    lib.sfk_hip_build_batch_device.restype = ctypes.c_int
    lib.sfk_hip_build.argtypes = [
        u8p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
        vp, ctypes.c_uint32, vp, ctypes.c_uint32, vp, ctypes.c_uint32,
        ctypes.POINTER(HipStats), ctypes.c_char_p, ctypes.c_size_t]
    lib.sfk_hip_build.restype = ctypes.c_int
    lib.sfk_hip_malloc.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(vp)]
    lib.sfk_hip_malloc.restype = ctypes.c_int
    lib.sfk_hip_free.argtypes = [vp]
    lib.sfk_hip_free.restype = ctypes.c_int
    lib.sfk_hip_memcpy_h2d.argtypes = [vp, vp, ctypes.c_uint64]
    lib.sfk_hip_memcpy_h2d.restype = ctypes.c_int
    lib.sfk_hip_memcpy_d2h.argtypes = [vp, vp, ctypes.c_uint64]
    lib.sfk_hip_memcpy_d2h.restype = ctypes.c_int
    _HIP = lib
    return lib


def _dtype(bits: int):
    return np.uint32 if bits == 32 else np.uint64


def hip_build(text: np.ndarray, construction_bits: int = 32, *, sa_bits: int | None = None,
              isa_bits: int | None = None, lcp_bits: int | None = None,
              retain_isa: bool = True, retain_lcp: bool = True, profile: bool = False,
              sampling_rate: int = 1):
    """Host-buffer build through sfk_hip_build.

    `text` is the logical text (encoded bytes 0..6 with the sentinel 0 last).
    With sampling_rate K > 1 the arrays keep the rows of text positions that
    are multiples of K (ceil(n / K) entries each).
    Returns (sa, isa or None, lcp or None, stats dict).
    """
    lib = load_hip()
    text = np.ascontiguousarray(text, dtype=np.uint8)
    n = int(text.size)
    narrow = n - 1 <= 0xFFFFFFFF
    sa_bits = sa_bits or (32 if narrow else 64)
    isa_bits = isa_bits or (32 if narrow else 64)
    lcp_bits = lcp_bits or (32 if narrow else 64)
    rows = (n + sampling_rate - 1) // sampling_rate if sampling_rate > 0 else n
    sa = np.empty(rows, dtype=_dtype(sa_bits))
    isa = np.empty(rows, dtype=_dtype(isa_bits)) if retain_isa else None
    lcp = np.empty(rows, dtype=_dtype(lcp_bits)) if retain_lcp else None
    flags = (SFK_HIP_RETAIN_ISA if retain_isa else 0) | (SFK_HIP_RETAIN_LCP if retain_lcp else 0)
    if profile:
        flags |= SFK_HIP_PROFILE_KERNELS
    stats = HipStats()
    err = ctypes.create_string_buffer(512)
    status = lib.sfk_hip_build(
        text.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), n, construction_bits, sampling_rate, flags,
        sa.ctypes.data, sa_bits,
        isa.ctypes.data if isa is not None else None, isa_bits,
        lcp.ctypes.data if lcp is not None else None, lcp_bits,
        ctypes.byref(stats), err, len(err))
    if status != 0:
        raise SufkitError(status, err.value.decode(errors="replace"))
    return sa, isa, lcp, stats.as_dict()


class DeviceBuilder:
    """A builder handle (one HIP stream + grow-only workspace)."""

    def __init__(self, device: int = 0):
        self.lib = load_hip()
        handle = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        status = self.lib.sfk_hip_builder_create(device, ctypes.byref(handle), err, len(err))
        if status != 0:
            raise SufkitError(status, err.value.decode(errors="replace"))
        self.handle = handle

    @property
    def stream(self) -> int:
        return int(self.lib.sfk_hip_builder_stream(self.handle) or 0)

    def build(self, d_text: int, n: int, construction_bits: int, d_sa: int, sa_bits: int,
              d_isa: int | None, isa_bits: int, d_lcp: int | None, lcp_bits: int,
              flags: int = SFK_HIP_RETAIN_ISA | SFK_HIP_RETAIN_LCP, sampling_rate: int = 1) -> dict:
        stats = HipStats()
        err = ctypes.create_string_buffer(512)
        status = self.lib.sfk_hip_build_device(
            self.handle, d_text, n, construction_bits, sampling_rate, flags, d_sa, sa_bits,
            d_isa, isa_bits, d_lcp, lcp_bits, ctypes.byref(stats), err, len(err))
        if status != 0:
            raise SufkitError(status, err.value.decode(errors="replace"))
        return stats.as_dict()

    def build_batch(self, d_text: int, n: int, offsets: np.ndarray, construction_bits: int,
                    d_sa: int, sa_bits: int, d_isa: int | None, isa_bits: int, d_lcp: int | None,
                    lcp_bits: int, flags: int = SFK_HIP_RETAIN_ISA | SFK_HIP_RETAIN_LCP) -> dict:
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        stats = HipStats()
        err = ctypes.create_string_buffer(512)
        status = self.lib.sfk_hip_build_batch_device(
            self.handle, d_text, n, offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
            offsets.size - 1, construction_bits, flags, d_sa, sa_bits, d_isa, isa_bits, d_lcp,
            lcp_bits, ctypes.byref(stats), err, len(err))
        if status != 0:
            raise SufkitError(status, err.value.decode(errors="replace"))
        return stats.as_dict()

    def close(self) -> None:
        if self.handle:
            self.lib.sfk_hip_builder_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_HOST = None


def host_library_path() -> Path:
    """libsufkit.so in the package, or the file SUFKIT_AMD_HOST_LIBRARY names
    (e.g. a sanitizer build of the same sources)."""
    override = os.environ.get("SUFKIT_AMD_HOST_LIBRARY")
    return Path(override) if override else LIB_DIR / "libsufkit.so"


def load_host() -> ctypes.CDLL:
    """Load libsufkit.so (host C++ library); raise if absent."""
    global _HOST
    if _HOST is not None:
        return _HOST
    path = host_library_path()
    if not path.exists():
        raise RuntimeError(f"{path} is missing: build it with `python -m sufkit_amd.build`")
    lib = ctypes.CDLL(str(path))
    lib.sfk_synth_plan.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double,
                                   ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
    lib.sfk_synth_plan.restype = ctypes.c_uint64
    lib.sfk_synth_generate.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double,
                                       ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint]
    lib.sfk_synth_generate.restype = ctypes.c_int
    _HOST = lib
    return lib


DEFAULT_SEED = 20260822


def synthetic_text(workload: int, seed: int = DEFAULT_SEED, ref: int = 0, scale: float = 1.0,
                   threads: int = 8) -> tuple[np.ndarray, np.ndarray]:
    """(logical text, contig lengths) of a benchmark workload (1..5)."""
    lib = load_host()
    count = ctypes.c_uint64()
    lens = np.zeros(4096, dtype=np.uint64)
    n = lib.sfk_synth_plan(workload, seed, ref, scale, lens.ctypes.data, lens.size, ctypes.byref(count))
    if n == 0:
        raise ValueError(f"unknown workload {workload}")
    text = np.empty(n, dtype=np.uint8)
    if lib.sfk_synth_generate(workload, seed, ref, scale, text.ctypes.data, n, threads) != 0:
        raise RuntimeError("synthetic generation failed")
    return text, lens[: count.value].copy()


class DeviceArray:
    """A device allocation through the C ABI (for tests and the benchmark)."""

    def __init__(self, nbytes: int, device: int = 0):
        self.lib = load_hip()
        ptr = ctypes.c_void_p()
        if self.lib.sfk_hip_malloc(device, max(int(nbytes), 1), ctypes.byref(ptr)) != 0:
            raise MemoryError(f"device allocation of {nbytes} bytes failed")
        self.ptr = int(ptr.value)
        self.nbytes = int(nbytes)

    def upload(self, array: np.ndarray) -> None:
        array = np.ascontiguousarray(array)
        assert array.nbytes <= self.nbytes
        if self.lib.sfk_hip_memcpy_h2d(self.ptr, array.ctypes.data, array.nbytes) != 0:
            raise RuntimeError("host to device copy failed")

    def download(self, dtype, count: int, offset_elems: int = 0) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if count and self.lib.sfk_hip_memcpy_d2h(
                out.ctypes.data, self.ptr + offset_elems * out.itemsize, out.nbytes) != 0:
            raise RuntimeError("device to host copy failed")
        return out

    def free(self) -> None:
        if self.ptr:
            self.lib.sfk_hip_free(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class CheckReport(ctypes.Structure):
    _fields_ = [(name, ctypes.c_uint64) for name in (
        "rows", "range_failures", "inverse_failures", "order_failures", "lcp_failures",
        "lcp_prefix_failures", "first_failed_row")] + [
        ("lcp_checked", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("seconds", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


def _check_lib():
    lib = load_hip()
    if not hasattr(lib, "_check_bound"):
        vp = ctypes.c_void_p
        lib.sfk_hip_digest.argtypes = [ctypes.c_int, vp, ctypes.c_uint64, ctypes.c_uint32,
                                       ctypes.POINTER(ctypes.c_uint64), ctypes.c_char_p, ctypes.c_size_t]
        lib.sfk_hip_digest.restype = ctypes.c_int
        lib.sfk_hip_check_index.argtypes = [ctypes.c_int, vp, ctypes.c_uint64, vp, ctypes.c_uint32, vp,
                                            ctypes.c_uint32, vp, ctypes.c_uint32, ctypes.c_uint64,
                                            ctypes.POINTER(CheckReport), ctypes.c_char_p, ctypes.c_size_t]
        lib.sfk_hip_check_index.restype = ctypes.c_int
        lib._check_bound = True
    return lib


def device_digest(ptr: int, count: int, value_bits: int, device: int = 0) -> int:
    """sfk_hip_digest of `count` device values (the oracle's sum digest)."""
    lib = _check_lib()
    out = ctypes.c_uint64()
    err = ctypes.create_string_buffer(512)
    status = lib.sfk_hip_digest(device, ptr, count, value_bits, ctypes.byref(out), err, len(err))
    if status != 0:
        raise SufkitError(status, err.value.decode(errors="replace"))
    return int(out.value)


def check_index(d_text: int, n: int, d_sa: int, sa_bits: int, d_isa: int, isa_bits: int,
                d_lcp: int | None = None, lcp_bits: int = 32, seed: int = 20260822,
                device: int = 0, raise_on_failure: bool = True) -> dict:
    """sfk_hip_check_index: order, permutation/inverse and LCP of a device
    index.  Returns the report; raises SufkitError on a failed check unless
    raise_on_failure is False (the report then carries status and message)."""
    lib = _check_lib()
    rep = CheckReport()
    err = ctypes.create_string_buffer(512)
    status = lib.sfk_hip_check_index(device, d_text, n, d_sa, sa_bits, d_isa, isa_bits, d_lcp, lcp_bits,
                                     seed, ctypes.byref(rep), err, len(err))
    out = rep.as_dict()
    out["status"] = status
    out["message"] = err.value.decode(errors="replace")
    if status != 0 and raise_on_failure:
        raise SufkitError(status, out["message"])
    return out


def hip_build_batch(texts: list[np.ndarray], construction_bits: int = 32, out_bits: int = 32,
                    builder: "DeviceBuilder | None" = None):
    """Batched build of several logical texts; returns [(sa, isa, lcp)] in
    each reference's own coordinates, and the stats."""
    offsets = np.zeros(len(texts) + 1, dtype=np.uint64)
    offsets[1:] = np.cumsum([t.size for t in texts])
    n = int(offsets[-1])
    joined = np.concatenate([np.asarray(t, np.uint8) for t in texts])
    d_text = DeviceArray(n + 64)
    d_text.upload(joined)
    w = out_bits // 8
    d_sa, d_isa, d_lcp = DeviceArray(n * w), DeviceArray(n * w), DeviceArray(n * w)
    own = builder is None
    builder = builder or DeviceBuilder(0)
    try:
        stats = builder.build_batch(d_text.ptr, n, offsets, construction_bits, d_sa.ptr, out_bits,
                                    d_isa.ptr, out_bits, d_lcp.ptr, out_bits)
        dt = _dtype(out_bits)
        sa, isa, lcp = d_sa.download(dt, n), d_isa.download(dt, n), d_lcp.download(dt, n)
    finally:
        if own:
            builder.close()
    parts = [(sa[a:b], isa[a:b], lcp[a:b]) for a, b in zip(offsets[:-1], offsets[1:])]
    return parts, stats


class ServeRequest(ctypes.Structure):
    _fields_ = [("arrival_seconds", ctypes.c_double), ("text", ctypes.c_void_p), ("n", ctypes.c_uint64)]


class ServeConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("slots", ctypes.c_uint32),
                ("queries_per_index", ctypes.c_uint64), ("min_query_len", ctypes.c_uint32),
                ("max_query_len", ctypes.c_uint32), ("query_seed", ctypes.c_uint64),
                ("keep_resident", ctypes.c_uint32), ("sample_queries", ctypes.c_uint32),
                ("digest_outputs", ctypes.c_uint32), ("fm_queries", ctypes.c_uint32),
                ("strict_budget", ctypes.c_uint32)]


class ServeResult(ctypes.Structure):
    _fields_ = [("arrival", ctypes.c_double), ("admitted", ctypes.c_double), ("built", ctypes.c_double),
                ("answered", ctypes.c_double), ("device_build_seconds", ctypes.c_double),
                ("n", ctypes.c_uint64), ("footprint_bytes", ctypes.c_uint64), ("count_sum", ctypes.c_uint64),
                ("count_hash", ctypes.c_uint64), ("output_hash", ctypes.c_uint64 * 3),
                ("status", ctypes.c_int), ("slot", ctypes.c_int)]


class ServeSummary(ctypes.Structure):
    _fields_ = [("makespan_seconds", ctypes.c_double), ("latency_p50", ctypes.c_double),
                ("latency_p99", ctypes.c_double), ("end_to_end_p50", ctypes.c_double),
                ("end_to_end_p99", ctypes.c_double), ("build_mbps", ctypes.c_double),
                ("queries_per_second", ctypes.c_double), ("peak_reserved_bytes", ctypes.c_uint64),
                ("requests_failed", ctypes.c_uint64), ("max_concurrent", ctypes.c_uint64),
                ("evicted", ctypes.c_uint64), ("first_error", ctypes.c_char * 256)]


def serve(texts: list[np.ndarray], arrivals: list[float], *, slots: int = 4,
          hbm_budget_bytes: int = 200 << 30, queries_per_index: int = 1_000_000,
          min_query_len: int = 20, max_query_len: int = 32, query_seed: int = 7,
          keep_resident: bool = True, sample_queries: int = 0, digest_outputs: bool = False,
          fm_queries: bool = True, strict_budget: bool = False):
    """Run the serving loop (sufkit_amd/host/serve.cpp) over host texts with
    the given arrival times; count queries by FM backward search (fm_queries)
    or SA binary search.  strict_budget: refuse (status 1, invalid input) a
    request whose footprint alone exceeds the budget instead of running it
    alone.  Returns (results, summary, samples)."""
    lib = load_host()
    lib.sfk_serve_run.argtypes = [ctypes.POINTER(ServeRequest), ctypes.c_uint64, ctypes.POINTER(ServeConfig),
                                  ctypes.POINTER(ServeResult), ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.POINTER(ServeSummary), ctypes.c_char_p,
                                  ctypes.c_size_t]
    lib.sfk_serve_run.restype = ctypes.c_int
    texts = [np.ascontiguousarray(t, dtype=np.uint8) for t in texts]
    count = len(texts)
    reqs = (ServeRequest * count)(*[ServeRequest(a, t.ctypes.data, t.size) for a, t in zip(arrivals, texts)])
    cfg = ServeConfig(hbm_budget_bytes, slots, queries_per_index, min_query_len, max_query_len,
                      query_seed, int(keep_resident), sample_queries, int(digest_outputs), int(fm_queries),
                      int(strict_budget))
    results = (ServeResult * count)()
    summary = ServeSummary()
    sp = np.zeros(max(1, count * sample_queries), np.uint64)
    sl = np.zeros(max(1, count * sample_queries), np.uint8)
    sc = np.zeros(max(1, count * sample_queries), np.uint32)
    err = ctypes.create_string_buffer(512)
    status = lib.sfk_serve_run(reqs, count, ctypes.byref(cfg), results, sp.ctypes.data, sl.ctypes.data,
                               sc.ctypes.data, ctypes.byref(summary), err, len(err))
    out = [{name: (list(getattr(r, name)) if name == "output_hash" else getattr(r, name))
            for name, _ in ServeResult._fields_} for r in results]
    summ = {name: getattr(summary, name) for name, _ in ServeSummary._fields_}
    if status != 0 and not any(r["status"] == 0 for r in out):
        raise SufkitError(status, err.value.decode(errors="replace") or "serving failed")
    samples = (sp.reshape(count, -1), sl.reshape(count, -1), sc.reshape(count, -1)) if sample_queries else None
    return out, summ, samples


def _prefix_lib():
    lib = load_hip()
    u64p = ctypes.POINTER(ctypes.c_uint64)
    lib.sfk_prefix_directory_select_k.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    lib.sfk_prefix_directory_select_k.restype = ctypes.c_uint32
    lib.sfk_prefix_directory_words.argtypes = [ctypes.c_uint32, ctypes.c_uint64]
    lib.sfk_prefix_directory_words.restype = ctypes.c_uint64
    lib.sfk_hip_prefix_directory.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                                             ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, u64p,
                                             ctypes.c_char_p, ctypes.c_size_t]
    lib.sfk_hip_prefix_directory.restype = ctypes.c_int
    lib.sfk_hip_build_fast.argtypes = [
        ctypes.POINTER(ctypes.c_uint8), ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
        ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32,
        ctypes.c_uint32, u64p, u64p, ctypes.POINTER(HipStats), ctypes.c_char_p, ctypes.c_size_t]
    lib.sfk_hip_build_fast.restype = ctypes.c_int
    lib.sfk_hip_index_prefix_k.argtypes = [ctypes.c_void_p]
    lib.sfk_hip_index_prefix_k.restype = ctypes.c_uint32
    return lib


def prefix_select_k(suffix_count: int, resident_core_bytes: int) -> int:
    """k of the Fast prefix directory for an index (0: none)."""
    return int(_prefix_lib().sfk_prefix_directory_select_k(suffix_count, resident_core_bytes))


def prefix_directory(text: np.ndarray, sa: np.ndarray, k: int):
    """Device prefix directory of a text and its SA (uint32 or uint64 rows):
    (begin, end, indexed k-mers, non-empty entries) as uint64 arrays of 4^k."""
    lib = _prefix_lib()
    text = np.ascontiguousarray(text, np.uint8)
    sa = np.ascontiguousarray(sa)
    n = text.size
    bits = 8 * sa.dtype.itemsize
    words = int(lib.sfk_prefix_directory_words(k, n))
    d_text, d_sa, d_table = DeviceArray(n + 64), DeviceArray(sa.nbytes), DeviceArray(words * 8)
    d_text.upload(text)
    d_sa.upload(sa)
    builder = DeviceBuilder(0)
    counts = (ctypes.c_uint64 * 2)()
    err = ctypes.create_string_buffer(512)
    try:
        status = lib.sfk_hip_prefix_directory(builder.handle, d_text.ptr, n, d_sa.ptr, bits, k, d_table.ptr,
                                              counts, err, len(err))
        if status != 0:
            raise SufkitError(status, err.value.decode(errors="replace"))
        table = d_table.download(np.uint64, words)
    finally:
        builder.close()
    return (*_split_directory(table, k, n), int(counts[0]), int(counts[1]))


def _split_directory(table: np.ndarray, k: int, n: int):
    entries = 1 << (2 * k)
    if n <= 0xFFFFFFFF:
        return table & np.uint64(0xFFFFFFFF), table >> np.uint64(32)
    return table[:entries], table[entries:]


def hip_build_fast(text: np.ndarray, prefix_k: int, construction_bits: int = 32):
    """sfk_hip_build_fast: SA/ISA/LCP (uint32) plus the prefix directory
    computed from the device SA: (sa, isa, lcp, begin, end, indexed,
    nonempty, stats)."""
    lib = _prefix_lib()
    text = np.ascontiguousarray(text, np.uint8)
    n = text.size
    sa, isa, lcp = np.empty(n, np.uint32), np.empty(n, np.uint32), np.empty(n, np.uint32)
    words = int(lib.sfk_prefix_directory_words(prefix_k, n))
    table = np.zeros(max(words, 1), np.uint64)
    counts = (ctypes.c_uint64 * 2)()
    stats = HipStats()
    err = ctypes.create_string_buffer(512)
    status = lib.sfk_hip_build_fast(text.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), n, construction_bits, 1,
                                    SFK_HIP_RETAIN_ISA | SFK_HIP_RETAIN_LCP, sa.ctypes.data, 32, isa.ctypes.data,
                                    32, lcp.ctypes.data, 32, prefix_k,
                                    table.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), counts,
                                    ctypes.byref(stats), err, len(err))
    if status != 0:
        raise SufkitError(status, err.value.decode(errors="replace"))
    begin, end = _split_directory(table[:words], prefix_k, n) if words else (None, None)
    return sa, isa, lcp, begin, end, int(counts[0]), int(counts[1]), stats.as_dict()


class TuneEntry(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("size_class", ctypes.c_uint32), ("block", ctypes.c_uint32),
                ("items", ctypes.c_uint32), ("us_per_pass", ctypes.c_double), ("chosen", ctypes.c_uint32),
                ("tuned", ctypes.c_uint32), ("digit_bits", ctypes.c_uint32)]


TUNE_KINDS = ("u32/u32", "u64/u32", "u32/u64", "u64/u64", "u32/u32 per reference", "u32/u64 per reference")


def _tune_lib():
    lib = load_hip()
    lib.sfk_hip_autotune.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(TuneEntry), ctypes.c_uint32,
                                     ctypes.POINTER(ctypes.c_uint32), ctypes.c_char_p, ctypes.c_size_t]
    lib.sfk_hip_autotune.restype = ctypes.c_int
    lib.sfk_hip_tune_get.argtypes = [ctypes.c_int, ctypes.POINTER(TuneEntry), ctypes.c_uint32,
                                     ctypes.POINTER(ctypes.c_uint32)]
    lib.sfk_hip_tune_get.restype = ctypes.c_int
    lib.sfk_hip_tune_set.argtypes = [ctypes.c_int, ctypes.POINTER(TuneEntry), ctypes.c_uint32, ctypes.c_char_p,
                                     ctypes.c_size_t]
    lib.sfk_hip_tune_set.restype = ctypes.c_int
    lib.sfk_hip_tune_reset.argtypes = [ctypes.c_int]
    lib.sfk_hip_tune_reset.restype = ctypes.c_int
    return lib


def _entries(arr, count):
    return [{name: getattr(arr[i], name) for name, _ in TuneEntry._fields_} for i in range(count)]


def autotune(max_elements: int, device: int = 0) -> list[dict]:
    """Time every compiled radix shape per sort kind and size class up to
    max_elements on the device, install the fastest, and return every
    measurement (chosen = 1 for the installed shapes)."""
    lib = _tune_lib()
    cap = 512
    arr = (TuneEntry * cap)()
    count = ctypes.c_uint32(0)
    err = ctypes.create_string_buffer(512)
    status = lib.sfk_hip_autotune(device, max_elements, arr, cap, ctypes.byref(count), err, len(err))
    if status != 0:
        raise SufkitError(status, err.value.decode(errors="replace") or "autotuning failed")
    return _entries(arr, min(count.value, cap))


def tune_get(device: int = 0) -> list[dict]:
    """The radix shape in use for every (kind, size class, digit width) of
    the device."""
    lib = _tune_lib()
    cap = 128
    arr = (TuneEntry * cap)()
    count = ctypes.c_uint32(0)
    status = lib.sfk_hip_tune_get(device, arr, cap, ctypes.byref(count))
    if status != 0:
        raise SufkitError(status, "tune_get failed")
    return _entries(arr, min(count.value, cap))


def tune_set(entries: list[dict], device: int = 0) -> None:
    """Install radix shapes (entries with chosen = 1), e.g. a saved table.
    digit_bits defaults to 8; us_per_pass, when given for both widths of a
    (kind, class), decides which width the sorts take."""
    lib = _tune_lib()
    arr = (TuneEntry * max(1, len(entries)))()
    for i, e in enumerate(entries):
        arr[i] = TuneEntry(e["kind"], e["size_class"], e["block"], e["items"], e.get("us_per_pass", 0.0),
                           e.get("chosen", 1), e.get("tuned", 1), e.get("digit_bits", 8))
    err = ctypes.create_string_buffer(512)
    status = lib.sfk_hip_tune_set(device, arr, len(entries), err, len(err))
    if status != 0:
        raise SufkitError(status, err.value.decode(errors="replace"))


def tune_reset(device: int = 0) -> None:
    """Back to the built-in radix shapes."""
    if _tune_lib().sfk_hip_tune_reset(device) != 0:
        raise SufkitError(2, "tune_reset failed")


def decode_pattern(code: int, length: int) -> str:
    """2-bit packed pattern (first symbol in the top bits) to an A/C/G/T string."""
    return "".join("ACGT"[(int(code) >> (62 - 2 * j)) & 3] for j in range(int(length)))


def encode_pattern(pattern: str) -> int:
    """A/C/G/T string (at most 32) to the 2-bit packed form of the count queries."""
    code = 0
    for j, ch in enumerate(pattern):
        code |= "ACGT".index(ch) << (62 - 2 * j)
    return code


def _index_lib():
    lib = load_hip()
    vp = ctypes.c_void_p
    lib.sfk_hip_index_create.argtypes = [ctypes.c_int, vp, ctypes.c_uint64, vp, ctypes.c_uint32,
                                         ctypes.POINTER(vp), ctypes.c_char_p, ctypes.c_size_t]
    lib.sfk_hip_index_create.restype = ctypes.c_int
    lib.sfk_hip_index_destroy.argtypes = [vp]
    lib.sfk_hip_index_destroy.restype = None
    lib.sfk_hip_index_build_fm.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
    lib.sfk_hip_index_build_fm.restype = ctypes.c_int
    for name in ("sfk_hip_index_count", "sfk_hip_index_fm_count"):
        getattr(lib, name).argtypes = [vp, vp, vp, ctypes.c_uint64, vp, vp]
        getattr(lib, name).restype = ctypes.c_int
    return lib


def index_count(text: np.ndarray, sa: np.ndarray, patterns: list[str], fm: bool = False) -> list[int]:
    """Count queries (at most 32 bases each) on a resident device index over
    a text and its SA: SA binary search, or FM backward search (fm)."""
    lib = _index_lib()
    text = np.ascontiguousarray(text, np.uint8)
    sa = np.ascontiguousarray(sa)
    n = text.size
    d_text, d_sa = DeviceArray(n + 64), DeviceArray(sa.nbytes)
    d_text.upload(text)
    d_sa.upload(sa)
    index = ctypes.c_void_p()
    err = ctypes.create_string_buffer(512)
    status = lib.sfk_hip_index_create(0, d_text.ptr, n, d_sa.ptr, 8 * sa.itemsize, ctypes.byref(index), err,
                                      len(err))
    if status != 0:
        raise SufkitError(status, err.value.decode(errors="replace"))
    try:
        if fm:
            status = lib.sfk_hip_index_build_fm(index, err, len(err))
            if status != 0:
                raise SufkitError(status, err.value.decode(errors="replace"))
        q = len(patterns)
        codes = np.array([encode_pattern(p) for p in patterns], np.uint64)
        lens = np.array([len(p) for p in patterns], np.uint8)
        d_p, d_l, d_c = DeviceArray(8 * q), DeviceArray(q), DeviceArray(4 * q)
        d_p.upload(codes)
        d_l.upload(lens)
        fn = lib.sfk_hip_index_fm_count if fm else lib.sfk_hip_index_count
        status = fn(index, d_p.ptr, d_l.ptr, q, d_c.ptr, None)
        if status != 0:
            raise SufkitError(status, "count failed")
        return [int(v) for v in d_c.download(np.uint32, q)]
    finally:
        lib.sfk_hip_index_destroy(index)


def device_crc32(data: np.ndarray, crc_in: int = 0, offset: int = 0) -> int:
    """zlib's crc32(crc_in, data) computed on the device: `data` is uploaded
    `offset` bytes into a fresh device buffer, so offset sets the alignment."""
    lib = load_hip()
    lib.sfk_hip_crc32.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                  ctypes.POINTER(ctypes.c_uint32), ctypes.c_char_p, ctypes.c_size_t]
    lib.sfk_hip_crc32.restype = ctypes.c_int
    data = np.ascontiguousarray(data, np.uint8)
    buf = DeviceArray(offset + data.size + 16)
    if data.size:
        if lib.sfk_hip_memcpy_h2d(buf.ptr + offset, data.ctypes.data, data.size) != 0:
            raise RuntimeError("host to device copy failed")
    out = ctypes.c_uint32()
    err = ctypes.create_string_buffer(512)
    status = lib.sfk_hip_crc32(0, buf.ptr + offset, data.size, crc_in, ctypes.byref(out), err, len(err))
    if status != 0:
        raise SufkitError(status, err.value.decode(errors="replace"))
    return int(out.value)

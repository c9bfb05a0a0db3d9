"""ctypes binding of the host library (sufkit_amd/lib/libsufkit.so).

Mirrors the reference's public surface for the construction path:
GenomeReference.from_records/from_fasta, SuffixArray.build/save/load,
IndexInfo, SuffixArray.count.  Errors raise SufkitError with the reference
ErrorCode name in `.code`.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import SufkitError, load_host

# SaBackend / CoordinateStorageWidth / SaAcceleration values (types.hpp:75-143)
BACKEND_AUTO, BACKEND_DIVSUFSORT, BACKEND_CAPS, BACKEND_HIP = 0, 1, 2, 3
ACCEL_NONE, ACCEL_LCP, ACCEL_LCP_CHILD, ACCEL_LCP_SUFFIX_LINK, ACCEL_FULL = 0, 1, 2, 3, 4


class _BuildOptions(ctypes.Structure):
    _fields_ = [("backend", ctypes.c_int), ("coordinate_width", ctypes.c_int),
                ("storage_width", ctypes.c_int), ("resource_profile", ctypes.c_int),
                ("acceleration", ctypes.c_int), ("threads", ctypes.c_uint32),
                ("sampling_rate", ctypes.c_uint32)]


class _BuildStats(ctypes.Structure):
    _fields_ = [(name, ctypes.c_double) for name in (
        "sa_seconds", "storage_compaction_seconds", "isa_seconds", "lcp_seconds",
        "text_prepare_seconds", "lcp_finalize_seconds", "total_seconds", "device_seconds",
        "transfer_seconds", "prefix_directory_seconds", "call_seconds")]


class _IndexInfo(ctypes.Structure):
    _fields_ = [("backend", ctypes.c_char * 32), ("backend_signature", ctypes.c_char * 96),
                ("format_version", ctypes.c_char * 8), ("coordinate_width", ctypes.c_uint8),
                ("stored_coordinate_width", ctypes.c_uint8), ("acceleration", ctypes.c_uint8),
                ("lcp_encoding", ctypes.c_uint8), ("sa_sampling_rate", ctypes.c_uint32)] + [
        (name, ctypes.c_uint64) for name in (
            "sequence_count", "total_bases", "text_symbols", "suffix_count", "ambiguous_bases",
            "fingerprint", "serialized_bytes", "text_bytes", "sa_bytes", "isa_bytes", "lcp_bytes",
            "resident_core_bytes", "auxiliary_bytes")] + [("prefix_directory_k", ctypes.c_uint32)]


_BOUND = False


def _lib():
    global _BOUND
    lib = load_host()
    if _BOUND:
        return lib
    vp, cp, sz = ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t
    u64p = ctypes.POINTER(ctypes.c_uint64)
    lib.sfk_reference_from_records.argtypes = [ctypes.POINTER(cp), ctypes.POINTER(cp), u64p,
                                               ctypes.c_uint64, ctypes.POINTER(vp), cp, sz]
    lib.sfk_reference_from_fasta.argtypes = [cp, ctypes.POINTER(vp), cp, sz]
    lib.sfk_reference_from_encoded.argtypes = [vp, u64p, ctypes.c_uint64, ctypes.POINTER(vp), cp, sz]
    lib.sfk_reference_free.argtypes = [vp]
    lib.sfk_reference_free.restype = None
    lib.sfk_reference_summary.argtypes = [vp, u64p]
    lib.sfk_reference_summary.restype = None
    lib.sfk_reference_text.argtypes = [vp, vp]
    lib.sfk_reference_text.restype = None
    lib.sfk_reference_contigs.argtypes = [vp, u64p]
    lib.sfk_reference_contigs.restype = None
    lib.sfk_sa_build.argtypes = [vp, ctypes.POINTER(_BuildOptions), ctypes.POINTER(vp),
                                 ctypes.POINTER(_BuildStats), cp, sz]
    lib.sfk_sa_save.argtypes = [vp, cp, ctypes.c_int, cp, sz]
    lib.sfk_sa_load.argtypes = [cp, ctypes.POINTER(vp), cp, sz]
    lib.sfk_sa_free.argtypes = [vp]
    lib.sfk_sa_free.restype = None
    lib.sfk_sa_info.argtypes = [vp, ctypes.POINTER(_IndexInfo)]
    lib.sfk_sa_info.restype = None
    for name in ("sfk_sa_copy_sa", "sfk_sa_copy_isa", "sfk_sa_copy_lcp"):
        getattr(lib, name).argtypes = [vp, u64p]
        getattr(lib, name).restype = ctypes.c_uint64
    lib.sfk_sa_count.argtypes = [vp, cp, ctypes.c_uint64, u64p, cp, sz]
    lib.sfk_hip_backend_available.restype = ctypes.c_int
    lib.sfk_sufidx_verify.argtypes = [cp, cp, sz]
    lib.sfk_fm_build.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp), cp, sz]
    lib.sfk_fm_free.argtypes = [vp]
    lib.sfk_fm_free.restype = None
    lib.sfk_fm_summary.argtypes = [vp, u64p]
    lib.sfk_fm_summary.restype = None
    lib.sfk_fm_equal_range.argtypes = [vp, cp, ctypes.c_uint64, u64p, cp, sz]
    lib.sfk_fm_equal_range_batch.argtypes = [vp, cp, u64p, ctypes.c_uint64, ctypes.c_uint32, u64p, cp, sz]
    lib.sfk_fm_count.argtypes = [vp, cp, ctypes.c_uint64, ctypes.c_int, u64p, cp, sz]
    lib.sfk_fm_count_batch.argtypes = [vp, cp, u64p, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint32, u64p, cp, sz]
    lib.sfk_fm_locate.argtypes = [vp, cp, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                  ctypes.POINTER(vp), cp, sz]
    lib.sfk_fm_hits_summary.argtypes = [vp, u64p]
    lib.sfk_fm_hits_summary.restype = None
    lib.sfk_fm_hits_get.argtypes = [vp, ctypes.POINTER(ctypes.c_uint32), u64p, u64p, ctypes.POINTER(ctypes.c_uint8)]
    lib.sfk_fm_hits_get.restype = None
    lib.sfk_fm_hits_free.argtypes = [vp]
    lib.sfk_fm_hits_free.restype = None
    lib.sfk_fm_save.argtypes = [vp, cp, ctypes.c_int, cp, sz]
    u32 = ctypes.c_uint32
    lib.sfk_save_device_index.argtypes = [vp, ctypes.c_int, vp, vp, u32, vp, u32, vp, u32, u32,
                                          ctypes.POINTER(_BuildOptions), cp, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_double), cp, sz]
    lib.sfk_save_device_index.restype = ctypes.c_int
    lib.sfk_file_crc32.argtypes = [cp, ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(u32), cp, sz]
    lib.sfk_file_crc32.restype = ctypes.c_int
    _BOUND = True
    return lib


def _check(status: int, err) -> None:
    if status != 0:
        raise SufkitError(status, err.value.decode(errors="replace"))


class GenomeReference:
    def __init__(self, handle):
        self._h = handle

    @classmethod
    def from_records(cls, records: list[tuple[str, bytes | str]]) -> "GenomeReference":
        lib = _lib()
        seqs = [s.encode() if isinstance(s, str) else bytes(s) for _, s in records]
        names = (ctypes.c_char_p * len(records))(*[n.encode() for n, _ in records])
        cseqs = (ctypes.c_char_p * len(records))(*seqs)
        lens = np.array([len(s) for s in seqs], dtype=np.uint64)
        out = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        _check(lib.sfk_reference_from_records(names, cseqs, lens.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                              len(records), ctypes.byref(out), err, len(err)), err)
        return cls(out)

    @classmethod
    def from_fasta(cls, path: str | Path) -> "GenomeReference":
        lib = _lib()
        out = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        _check(lib.sfk_reference_from_fasta(str(path).encode(), ctypes.byref(out), err, len(err)), err)
        return cls(out)

    @classmethod
    def from_encoded(cls, text: np.ndarray, contig_lengths: np.ndarray) -> "GenomeReference":
        """From a logical text as produced by the synthetic generator."""
        lib = _lib()
        lengths = np.ascontiguousarray(contig_lengths, dtype=np.uint64)
        bases = np.ascontiguousarray(text[:-1][text[:-1] != 1], dtype=np.uint8)
        out = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        _check(lib.sfk_reference_from_encoded(bases.ctypes.data, lengths.ctypes.data_as(
            ctypes.POINTER(ctypes.c_uint64)), lengths.size, ctypes.byref(out), err, len(err)), err)
        return cls(out)

    def summary(self) -> dict:
        out = (ctypes.c_uint64 * 5)()
        _lib().sfk_reference_summary(self._h, out)
        return dict(sequence_count=out[0], total_bases=out[1], ambiguous_bases=out[2],
                    fingerprint=out[3], text_symbols=out[4])

    def text(self) -> np.ndarray:
        buf = np.empty(self.summary()["text_symbols"], dtype=np.uint8)
        _lib().sfk_reference_text(self._h, buf.ctypes.data)
        return buf

    def contigs(self) -> np.ndarray:
        count = self.summary()["sequence_count"]
        out = np.empty(3 * count, dtype=np.uint64)
        _lib().sfk_reference_contigs(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        return out.reshape(count, 3)

    def __del__(self):
        if getattr(self, "_h", None):
            _lib().sfk_reference_free(self._h)
            self._h = None


@dataclass
class BuildOptions:
    backend: int = BACKEND_AUTO
    coordinate_width: int = 0
    storage_width: int = 0
    resource_profile: int = 0
    acceleration: int = ACCEL_LCP_SUFFIX_LINK
    threads: int = 1
    sampling_rate: int = 1


class SuffixArray:
    def __init__(self, handle, stats: dict | None = None):
        self._h = handle
        self.stats = stats or {}

    @classmethod
    def build(cls, reference: GenomeReference, options: BuildOptions | None = None) -> "SuffixArray":
        lib = _lib()
        o = options or BuildOptions()
        copts = _BuildOptions(o.backend, o.coordinate_width, o.storage_width, o.resource_profile,
                              o.acceleration, o.threads, o.sampling_rate)
        stats = _BuildStats()
        out = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        _check(lib.sfk_sa_build(reference._h, ctypes.byref(copts), ctypes.byref(out), ctypes.byref(stats),
                                err, len(err)), err)
        return cls(out, {name: getattr(stats, name) for name, _ in _BuildStats._fields_})

    @classmethod
    def load(cls, path: str | Path) -> "SuffixArray":
        lib = _lib()
        out = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        _check(lib.sfk_sa_load(str(path).encode(), ctypes.byref(out), err, len(err)), err)
        return cls(out)

    def save(self, path: str | Path, overwrite: bool = False) -> None:
        err = ctypes.create_string_buffer(512)
        _check(_lib().sfk_sa_save(self._h, str(path).encode(), int(overwrite), err, len(err)), err)

    def info(self) -> dict:
        info = _IndexInfo()
        _lib().sfk_sa_info(self._h, ctypes.byref(info))
        out = {name: getattr(info, name) for name, _ in _IndexInfo._fields_}
        for key in ("backend", "backend_signature", "format_version"):
            out[key] = out[key].decode()
        return out

    def _copy(self, fn) -> np.ndarray | None:
        lib = _lib()
        n = getattr(lib, fn)(self._h, None)
        if n == 0:
            return None
        out = np.empty(n, dtype=np.uint64)
        getattr(lib, fn)(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        return out

    def sa(self) -> np.ndarray:
        return self._copy("sfk_sa_copy_sa")

    def isa(self) -> np.ndarray | None:
        return self._copy("sfk_sa_copy_isa")

    def lcp(self) -> np.ndarray | None:
        return self._copy("sfk_sa_copy_lcp")

    def count(self, pattern: str) -> int:
        out = ctypes.c_uint64()
        err = ctypes.create_string_buffer(512)
        p = pattern.encode()
        _check(_lib().sfk_sa_count(self._h, p, len(p), ctypes.byref(out), err, len(err)), err)
        return int(out.value)

    def __del__(self):
        if getattr(self, "_h", None):
            _lib().sfk_sa_free(self._h)
            self._h = None


# FmBackend / StrandMode / Strand values (types.hpp:200-229).
FM_SDSL_CSA_WT_HUFF, FM_SDSL_CSA_WT_BALANCED, FM_SDSL_CSA_SADA, FM_SDSL_CSA_WT_EPR = 1, 2, 3, 4
FORWARD, REVERSE_COMPLEMENT, BOTH = 0, 1, 2


def _joined(patterns: list[str]):
    data = [p.encode() for p in patterns]
    offsets = np.zeros(len(data) + 1, dtype=np.uint64)
    if data:
        offsets[1:] = np.cumsum([len(d) for d in data])
    return b"".join(data) or b"\0", offsets


class FmIndex:
    """FmIndex over the device FM-index built from the GPU SA (fm_index.hpp)."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def build(cls, reference: GenomeReference, backend: int = FM_SDSL_CSA_WT_HUFF) -> "FmIndex":
        out = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        _check(_lib().sfk_fm_build(reference._h, backend, ctypes.byref(out), err, len(err)), err)
        return cls(out)

    def summary(self) -> dict:
        out = (ctypes.c_uint64 * 7)()
        _lib().sfk_fm_summary(self._h, out)
        keys = ("text_symbols", "sequence_count", "total_bases", "fingerprint", "fm_device_bytes",
                "prefix_directory_k", "sa_bits")
        return dict(zip(keys, (int(v) for v in out)))

    def equal_range(self, pattern: str) -> tuple[int, int]:
        out = (ctypes.c_uint64 * 2)()
        err = ctypes.create_string_buffer(512)
        p = pattern.encode()
        _check(_lib().sfk_fm_equal_range(self._h, p, len(p), out, err, len(err)), err)
        return int(out[0]), int(out[1])

    def equal_range_batch(self, patterns: list[str], batch_width: int = 0) -> list[tuple[int, int]]:
        data, offsets = _joined(patterns)
        out = np.zeros(2 * max(len(patterns), 1), dtype=np.uint64)
        err = ctypes.create_string_buffer(512)
        _check(_lib().sfk_fm_equal_range_batch(self._h, data, offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                               len(patterns), batch_width,
                                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), err, len(err)),
               err)
        return [(int(out[2 * i]), int(out[2 * i + 1])) for i in range(len(patterns))]

    def count(self, pattern: str, strands: int = FORWARD) -> int:
        out = ctypes.c_uint64()
        err = ctypes.create_string_buffer(512)
        p = pattern.encode()
        _check(_lib().sfk_fm_count(self._h, p, len(p), strands, ctypes.byref(out), err, len(err)), err)
        return int(out.value)

    def count_batch(self, patterns: list[str], strands: int = FORWARD, batch_width: int = 0) -> list[int]:
        data, offsets = _joined(patterns)
        out = np.zeros(max(len(patterns), 1), dtype=np.uint64)
        err = ctypes.create_string_buffer(512)
        _check(_lib().sfk_fm_count_batch(self._h, data, offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                         len(patterns), strands, batch_width,
                                         out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), err, len(err)), err)
        return [int(v) for v in out[:len(patterns)]]

    def locate(self, pattern: str, strands: int = FORWARD, max_hits: int | None = None) -> dict:
        """{"total_hits", "hits": [(sequence_id, position, length, strand)], "truncated"}."""
        lib = _lib()
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        p = pattern.encode()
        _check(lib.sfk_fm_locate(self._h, p, len(p), strands, int(max_hits is not None), max_hits or 0,
                                 ctypes.byref(h), err, len(err)), err)
        try:
            summary = (ctypes.c_uint64 * 3)()
            lib.sfk_fm_hits_summary(h, summary)
            size = int(summary[1])
            ids = np.zeros(max(size, 1), dtype=np.uint32)
            pos = np.zeros(max(size, 1), dtype=np.uint64)
            lens = np.zeros(max(size, 1), dtype=np.uint64)
            strs = np.zeros(max(size, 1), dtype=np.uint8)
            lib.sfk_fm_hits_get(h, ids.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                pos.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                lens.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                strs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
        finally:
            lib.sfk_fm_hits_free(h)
        hits = [(int(ids[i]), int(pos[i]), int(lens[i]), int(strs[i])) for i in range(size)]
        return {"total_hits": int(summary[0]), "hits": hits, "truncated": bool(summary[2])}

    def save(self, path: str | Path, overwrite: bool = False) -> None:
        err = ctypes.create_string_buffer(512)
        _check(_lib().sfk_fm_save(self._h, str(path).encode(), int(overwrite), err, len(err)), err)

    def __del__(self):
        if getattr(self, "_h", None):
            _lib().sfk_fm_free(self._h)
            self._h = None


def verify_sufidx(path: str | Path) -> None:
    err = ctypes.create_string_buffer(512)
    _check(_lib().sfk_sufidx_verify(str(path).encode(), err, len(err)), err)


def resolve_build_policy(backend: int, coordinate_width: int, storage_width: int, profile: int,
                         text_symbols: int, threads: int = 1) -> dict:
    lib = _lib()
    lib.sfk_resolve_build_policy.argtypes = [ctypes.c_int] * 4 + [ctypes.c_uint64, ctypes.c_uint32,
                                             ctypes.POINTER(ctypes.c_int), ctypes.c_char_p, ctypes.c_size_t]
    out = (ctypes.c_int * 5)()
    err = ctypes.create_string_buffer(512)
    _check(lib.sfk_resolve_build_policy(backend, coordinate_width, storage_width, profile, text_symbols,
                                        threads, out, err, len(err)), err)
    return dict(backend=out[0], construction_width=out[1], stored_width=out[2], isa_width=out[3],
                lcp_width=out[4])


def hip_backend_available() -> bool:
    return bool(_lib().sfk_hip_backend_available())


def file_crc32(path: str | Path, offset: int = 0, size: int | None = None) -> int:
    """zlib CRC32 of a byte range of a file, as the container reader checks
    section checksums (several threads, joined)."""
    if size is None:
        size = Path(path).stat().st_size - offset
    out = ctypes.c_uint32()
    err = ctypes.create_string_buffer(512)
    _check(_lib().sfk_file_crc32(str(path).encode(), offset, size, ctypes.byref(out), err, len(err)), err)
    return int(out.value)


def save_device_index(reference: GenomeReference, path: str | Path, options: BuildOptions | None = None, *,
                      construction_bits: int = 32, sa_bits: int | None = None, isa_bits: int | None = None,
                      lcp_bits: int = 32, overwrite: bool = False, device_sa: np.ndarray | None = None) -> dict:
    """Build `reference` on the device (sfk_hip_build_device) and write its
    .sufidx straight from the device arrays (SaveDeviceIndex).  The device
    widths default to 64 bits when the stored width is wider than 32.
    `device_sa` replaces the built SA by the given values (error-path tests).
    Returns the save statistics."""
    from . import DeviceArray, DeviceBuilder, SFK_HIP_RETAIN_ISA, SFK_HIP_RETAIN_LCP
    o = options or BuildOptions()
    wide = o.storage_width in (40, 48, 64)
    sa_bits = sa_bits or (64 if wide else 32)
    isa_bits = isa_bits or (64 if wide else 32)
    text = reference.text()
    if text.size == 0 or text[-1] != 0:
        text = np.append(text, np.uint8(0))
    n = int(text.size)
    low_memory = o.resource_profile == 1  # byte-coded LCP made from the device ISA and raw LCP
    keep_isa = low_memory or o.acceleration == ACCEL_LCP_SUFFIX_LINK
    keep_lcp = low_memory or o.acceleration != ACCEL_NONE
    d_text = DeviceArray(n + 64)
    d_text.upload(text)
    d_sa = DeviceArray(n * sa_bits // 8)
    d_isa = DeviceArray(n * isa_bits // 8)
    d_lcp = DeviceArray(n * lcp_bits // 8)
    flags = (SFK_HIP_RETAIN_ISA if keep_isa else 0) | (SFK_HIP_RETAIN_LCP if keep_lcp else 0)
    with DeviceBuilder() as builder:
        builder.build(d_text.ptr, n, construction_bits, d_sa.ptr, sa_bits, d_isa.ptr if keep_isa else None,
                      isa_bits, d_lcp.ptr if keep_lcp else None, lcp_bits, flags=flags)
    if device_sa is not None:
        d_sa.upload(np.ascontiguousarray(device_sa, dtype=np.uint64 if sa_bits == 64 else np.uint32))
    return save_device_arrays(reference, path, d_text.ptr, d_sa.ptr, sa_bits, d_isa.ptr if keep_isa else None,
                              isa_bits, d_lcp.ptr if keep_lcp else None, lcp_bits, construction_bits, o,
                              overwrite=overwrite)


def save_device_arrays(reference: GenomeReference, path: str | Path, d_text: int, d_sa: int, sa_bits: int,
                       d_isa: int | None, isa_bits: int, d_lcp: int | None, lcp_bits: int,
                       construction_bits: int, options: BuildOptions | None = None, *, overwrite: bool = False,
                       device: int = 0) -> dict:
    """SaveDeviceIndex over device arrays already built (device pointers)."""
    o = options or BuildOptions()
    copts = _BuildOptions(o.backend, o.coordinate_width, o.storage_width, o.resource_profile,
                          o.acceleration, o.threads, o.sampling_rate)
    stats = (ctypes.c_double * 4)()
    err = ctypes.create_string_buffer(512)
    _check(_lib().sfk_save_device_index(reference._h, device, d_text, d_sa, sa_bits, d_isa, isa_bits, d_lcp,
                                        lcp_bits, construction_bits, ctypes.byref(copts), str(path).encode(),
                                        int(overwrite), stats, err, len(err)), err)
    return dict(seconds=stats[0], write_seconds=stats[1], validate_seconds=stats[2], bytes=int(stats[3]))

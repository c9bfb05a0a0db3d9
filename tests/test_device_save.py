"""Streaming .sufidx writer from device buffers (SaveDeviceIndex) and the
device CRC-32 behind it.

The file written from the device arrays must be the file the host path
(SuffixArray::Build + Save) writes, byte for byte, and hence equal to the
reference's own Save output apart from the backend byte and the header CRC
(the golden files under tests/golden/sufidx were written by the compiled
reference).  The device CRC must equal zlib's crc32 for every length,
alignment and starting value."""
import json
import os
import struct
import zlib
from pathlib import Path

import numpy as np
import pytest

from sufkit_amd import SufkitError, api
from tests.golden.make_golden import golden_references
from tests.test_host_gpu import parse

GOLDEN = Path(__file__).resolve().parent / "golden"


def contigs(rng, total: int, count: int) -> list[bytes]:
    """`count` contigs of A/C/G/T totalling `total` bases, with N runs and a
    copied repeat in each."""
    letters = np.frombuffer(b"ACGT", np.uint8)
    out = []
    for c in range(count):
        seq = letters[rng.integers(0, 4, total // count)]
        for _ in range(5):
            at = int(rng.integers(0, seq.size - 5000))
            seq[at:at + int(rng.integers(10, 4000))] = ord("N")
        src, dst = sorted(int(x) for x in rng.integers(0, seq.size - 20000, 2))
        seq[dst:dst + 15000] = seq[src:src + 15000].copy()
        out.append(seq.tobytes())
    return out


def golden_items():
    return sorted(json.loads((GOLDEN / "sufidx_index.json").read_text()).items())


# ----------------------------------------------------------------- CPU ---

def test_file_crc_matches_zlib_over_many_pieces(tmp_path):
    """The container reader checksums sections in pieces on several threads
    and joins them; the join must equal one zlib pass."""
    rng = np.random.default_rng(3)
    data = rng.integers(0, 256, (64 << 20) * 2 + 12345, dtype=np.uint8).tobytes()
    path = tmp_path / "blob"
    path.write_bytes(data)
    for offset, size in [(0, len(data)), (7, len(data) - 7), (100, 64 << 20), (5, (16 << 20) + 1), (11, (16 << 20) - 1), (3, 0),
                         (len(data) - 9, 9)]:
        assert api.file_crc32(path, offset, size) == zlib.crc32(data[offset:offset + size]), (offset, size)


def test_file_crc_reports_a_short_file(tmp_path):
    path = tmp_path / "blob"
    path.write_bytes(b"abc")
    with pytest.raises(SufkitError):
        api.file_crc32(path, 0, 10)


# ----------------------------------------------------------------- GPU ---

@pytest.mark.gpu
@pytest.mark.parametrize("offset", [0, 1, 3, 15])
def test_device_crc_equals_zlib(hip, offset):
    rng = np.random.default_rng(offset)
    blob = rng.integers(0, 256, 3 * 65536 + 70000, dtype=np.uint8)
    sizes = [0, 1, 3, 15, 16, 17, 255, 256, 4096, 65535, 65536, 65537, 131072 + 5, 3 * 65536 + 999, blob.size]
    for size in sizes:
        for crc_in in (0, 0xDEADBEEF):
            data = blob[:size]
            got = hip.device_crc32(data, crc_in, offset)
            assert got == zlib.crc32(data.tobytes(), crc_in), (size, crc_in, offset)


@pytest.mark.gpu
def test_device_crc_of_a_large_buffer(hip):
    rng = np.random.default_rng(9)
    data = rng.integers(0, 256, (40 << 20) + 77, dtype=np.uint8)
    assert hip.device_crc32(data) == zlib.crc32(data.tobytes())
    # run-heavy bytes too (every table entry path of a constant stream)
    runs = np.zeros(5 << 20, np.uint8)
    assert hip.device_crc32(runs, 0x12345678) == zlib.crc32(runs.tobytes(), 0x12345678)


@pytest.mark.gpu
@pytest.mark.parametrize("fname,meta", golden_items())
def test_device_save_matches_the_reference_save(tmp_path, hip, fname, meta):
    records = golden_references()[meta["reference"]]
    ref = api.GenomeReference.from_records(records)
    opts = api.BuildOptions(backend=api.BACKEND_HIP, coordinate_width=meta["coordinate_width"],
                            storage_width=meta["storage_width"], acceleration=meta["acceleration"],
                            resource_profile=meta.get("resource_profile", 0))
    bits = meta["coordinate_width"] or 32
    out = tmp_path / "device.sufidx"
    stats = api.save_device_index(ref, out, opts, construction_bits=bits)
    ours = out.read_bytes()
    assert stats["bytes"] == len(ours)
    theirs = (GOLDEN / "sufidx" / fname).read_bytes()
    h_ours, s_ours = parse(ours)
    h_theirs, s_theirs = parse(theirs)
    assert s_ours.keys() == s_theirs.keys()
    for stype in s_theirs:
        assert s_ours[stype] == s_theirs[stype], f"section {stype} differs"
    assert len(ours) == len(theirs)
    diff = [i for i in range(len(h_ours)) if h_ours[i] != h_theirs[i]]
    assert set(diff) <= {14, 72, 73, 74, 75}, diff
    # and equal to the host path's file, header included
    host = tmp_path / "host.sufidx"
    api.SuffixArray.build(ref, opts).save(host)
    assert host.read_bytes() == ours


@pytest.mark.gpu
@pytest.mark.parametrize("storage,sa_bits,isa_bits,lcp_bits,construction", [
    (0, 32, 32, 32, 32),
    (0, 64, 64, 64, 64),   # 64-bit device arrays stored as native32 / raw32
    (40, 64, 64, 32, 64),
    (48, 64, 32, 64, 32),  # 32-bit device ISA widened to the split planes
    (64, 64, 64, 32, 64),
])
def test_device_save_of_a_multi_chunk_reference(tmp_path, hip, storage, sa_bits, isa_bits, lcp_bits,
                                                construction):
    """A reference whose SA plane spans several 64 MiB chunks (and several
    CRC pieces on reread): the device file equals the host path's file and
    loads back to the same arrays."""
    rng = np.random.default_rng(storage + sa_bits)
    ref = api.GenomeReference.from_records([(f"c{i}", s) for i, s in enumerate(contigs(rng, 17_500_000, 5))])
    opts = api.BuildOptions(backend=api.BACKEND_HIP, coordinate_width=construction, storage_width=storage)
    dev = tmp_path / "device.sufidx"
    api.save_device_index(ref, dev, opts, construction_bits=construction, sa_bits=sa_bits, isa_bits=isa_bits,
                          lcp_bits=lcp_bits)
    host_index = api.SuffixArray.build(ref, opts)
    host = tmp_path / "host.sufidx"
    host_index.save(host)
    assert os.path.getsize(dev) == os.path.getsize(host)
    assert api.file_crc32(dev) == api.file_crc32(host)
    assert dev.read_bytes() == host.read_bytes()
    back = api.SuffixArray.load(dev)
    assert np.array_equal(back.sa(), host_index.sa())
    assert np.array_equal(back.lcp(), host_index.lcp())


@pytest.mark.gpu
def test_device_save_without_isa_or_lcp(tmp_path, hip):
    ref = api.GenomeReference.from_records([("a", "ACGTNNACGTTTGA" * 300), ("b", "GATTACA" * 500)])
    for accel in (api.ACCEL_NONE, api.ACCEL_LCP):
        opts = api.BuildOptions(backend=api.BACKEND_HIP, acceleration=accel)
        dev, host = tmp_path / f"d{accel}.sufidx", tmp_path / f"h{accel}.sufidx"
        api.save_device_index(ref, dev, opts)
        api.SuffixArray.build(ref, opts).save(host)
        assert dev.read_bytes() == host.read_bytes()


@pytest.mark.gpu
def test_device_save_rejects_values_outside_the_stored_width(tmp_path, hip):
    """A device SA value that native32 cannot hold is an error, and no file
    is published."""
    ref = api.GenomeReference.from_records([("a", "ACGT" * 40)])
    n = 162  # 160 bases, the contig separator, the sentinel
    bad = np.arange(n, dtype=np.uint64)
    bad[7] = 1 << 33
    out = tmp_path / "bad.sufidx"
    with pytest.raises(SufkitError) as e:
        api.save_device_index(ref, out, api.BuildOptions(backend=api.BACKEND_HIP), sa_bits=64, device_sa=bad)
    assert e.value.status == 1  # kInvalidInput
    assert not out.exists()
    assert list(tmp_path.iterdir()) == []


@pytest.mark.gpu
def test_device_save_keeps_an_existing_file(tmp_path, hip):
    ref = api.GenomeReference.from_records([("a", "ACGT" * 40)])
    out = tmp_path / "x.sufidx"
    out.write_bytes(b"keep")
    with pytest.raises(SufkitError) as e:
        api.save_device_index(ref, out, api.BuildOptions(backend=api.BACKEND_HIP))
    assert e.value.status == 2  # kIoError
    assert out.read_bytes() == b"keep"
    api.save_device_index(ref, out, api.BuildOptions(backend=api.BACKEND_HIP), overwrite=True)
    assert np.array_equal(api.SuffixArray.load(out).sa(),
                          api.SuffixArray.build(ref, api.BuildOptions(backend=api.BACKEND_HIP)).sa())


@pytest.mark.gpu
def test_device_save_rejects_sampling(tmp_path, hip):
    ref = api.GenomeReference.from_records([("a", "ACGT" * 40)])
    with pytest.raises(SufkitError) as e:
        api.save_device_index(ref, tmp_path / "s.sufidx", api.BuildOptions(backend=api.BACKEND_HIP, sampling_rate=2))
    assert e.value.status == 1


@pytest.mark.gpu
def test_a_corrupted_section_is_detected_on_load(tmp_path, hip):
    """Flip one byte deep inside a multi-piece SA section of a device-written
    file: the reread checksum must catch it."""
    rng = np.random.default_rng(4)
    ref = api.GenomeReference.from_records([(f"c{i}", s) for i, s in enumerate(contigs(rng, 18_000_000, 3))])
    out = tmp_path / "c.sufidx"
    api.save_device_index(ref, out, api.BuildOptions(backend=api.BACKEND_HIP))
    data = bytearray(out.read_bytes())
    count = struct.unpack_from("<I", data, 24)[0]
    for i in range(count):
        stype, _, offset, size, _, _ = struct.unpack_from("<IIQQII", data, 80 + 32 * i)
        if stype == 3:  # SA
            at = offset + size - (3 << 20)
    data[at] ^= 0x40
    out.write_bytes(bytes(data))
    with pytest.raises(SufkitError) as e:
        api.SuffixArray.load(out)
    assert e.value.status == 4  # kCorruptIndex


@pytest.mark.gpu
@pytest.mark.parametrize("construction,storage", [(32, 0), (64, 40)])
def test_device_save_of_a_low_memory_index(tmp_path, hip, construction, storage):
    """Low-memory: byte-coded LCP (capped bytes streamed from the device raw
    LCP, anchors found on the device), no ISA section; the file equals the
    host path's on a multi-chunk reference whose repeats need overflow
    anchors."""
    rng = np.random.default_rng(7 + construction)
    seqs = contigs(rng, 17_500_000, 4)
    unit = seqs[0][1000:4000]
    seqs = [s + unit * 3 + b"A" * 900 for s in seqs]
    ref = api.GenomeReference.from_records([(f"c{i}", s) for i, s in enumerate(seqs)])
    opts = api.BuildOptions(backend=api.BACKEND_HIP, coordinate_width=construction, storage_width=storage,
                            resource_profile=1)
    dev, host = tmp_path / "device.sufidx", tmp_path / "host.sufidx"
    api.save_device_index(ref, dev, opts, construction_bits=construction)
    index = api.SuffixArray.build(ref, opts)
    index.save(host)
    assert dev.read_bytes() == host.read_bytes()
    back = api.SuffixArray.load(dev)
    assert back.info()["lcp_encoding"] == 2 and back.isa() is None
    assert np.array_equal(back.lcp(), index.lcp())


@pytest.mark.gpu
@pytest.mark.parametrize("fault", ["lcp0", "anchor", "isa"])
def test_low_memory_device_save_rejects_bad_arrays(tmp_path, hip, fault):
    """The Low-memory save checks what BuildByteCoded checks before coding
    (lcp_storage.cpp:237-297): LCP[0] == 0, an ISA inverse to the SA, anchor
    values within the anchor coordinates; a failure throws kBuildFailure and
    publishes no file.  The arrays are the oracle's, altered on the CPU."""
    import oracle
    from sufkit_amd import DeviceArray

    ref = api.GenomeReference.from_records([("a", "ACGTTGCA" * 300 + "A" * 700), ("b", "TTGACCA" * 200)])
    text = ref.text()
    n = int(text.size)
    sa, isa, lcp = oracle.fast_build(text)
    lcp = lcp.astype(np.uint64)
    if fault == "lcp0":
        lcp[0] = 3
    elif fault == "anchor":
        # a run of overflow values beyond 32-bit anchors (n < 2^32)
        lcp[int(isa[5])] = (1 << 33) + 7
    else:
        isa = isa.copy()
        isa[[1, 2]] = isa[[2, 1]]
    arrays = [DeviceArray(n + 64)] + [DeviceArray(n * 8) for _ in range(3)]
    arrays[0].upload(text)
    for d, a in zip(arrays[1:], (sa, isa, lcp)):
        d.upload(np.ascontiguousarray(a, dtype=np.uint64))
    out = tmp_path / "bad.sufidx"
    opts = api.BuildOptions(backend=api.BACKEND_HIP, resource_profile=1)
    with pytest.raises(SufkitError) as e:
        api.save_device_arrays(ref, out, arrays[0].ptr, arrays[1].ptr, 64, arrays[2].ptr, 64, arrays[3].ptr, 64,
                               32, opts)
    assert e.value.status == 6  # kBuildFailure
    assert not out.exists()

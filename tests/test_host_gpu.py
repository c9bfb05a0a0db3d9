"""SuffixArray::Build with the HIP backend through the host library, and the
.sufidx output compared byte for byte with the reference's own Save.

With a new stored backend ID the file differs from the divsufsort file only
at header byte 14 (backend) and the header CRC (offset 72); every section
payload and section CRC must be identical."""
import json
import struct
from pathlib import Path

import numpy as np
import pytest

import oracle
from sufkit_amd import SufkitError, api
from tests import synth
from tests.golden.make_golden import golden_references

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def parse(data: bytes):
    count, header_size = struct.unpack_from("<II", data, 24)
    sections = {}
    for i in range(count):
        stype, flags, offset, size, crc, _ = struct.unpack_from("<IIQQII", data, 80 + 32 * i)
        sections[stype] = (flags, crc, data[offset:offset + size])
    return data[:header_size], sections


def golden_items():
    return sorted(json.loads((GOLDEN / "sufidx_index.json").read_text()).items())


@pytest.mark.parametrize("fname,meta", golden_items())
def test_saved_file_matches_reference_save(tmp_path, hip, fname, meta):
    records = golden_references()[meta["reference"]]
    ref = api.GenomeReference.from_records(records)
    opts = api.BuildOptions(backend=api.BACKEND_HIP, coordinate_width=meta["coordinate_width"],
                            storage_width=meta["storage_width"], acceleration=meta["acceleration"],
                            resource_profile=meta.get("resource_profile", 0))
    sa = api.SuffixArray.build(ref, opts)
    out = tmp_path / "ours.sufidx"
    sa.save(out)
    ours = out.read_bytes()
    theirs = (GOLDEN / "sufidx" / fname).read_bytes()
    h_ours, s_ours = parse(ours)
    h_theirs, s_theirs = parse(theirs)
    assert s_ours.keys() == s_theirs.keys()
    for stype in s_theirs:
        assert s_ours[stype] == s_theirs[stype], f"section {stype} differs"
    assert len(ours) == len(theirs)
    diff = [i for i in range(len(h_ours)) if h_ours[i] != h_theirs[i]]
    assert set(diff) <= {14, 72, 73, 74, 75}, diff
    # hip32/hip64 in place of divsufsort32/64
    assert h_theirs[14] in (1, 2) and h_ours[14] == h_theirs[14] + 4
    assert h_ours[15] == h_theirs[15]  # construction width
    # our loader accepts our file
    back = api.SuffixArray.load(out)
    assert np.array_equal(back.sa(), sa.sa())
    api.verify_sufidx(out)


def test_build_fixture_reference(hip, legacy_fixtures):
    ref = api.GenomeReference.from_records([("chr1", b"ACGTNACGT"), ("chr2", b"TTTACGTAAA"),
                                            ("chr3", b"AAA")])
    sa = api.SuffixArray.build(ref)
    v = legacy_fixtures["legacy-sa-v1.1"]["vectors"]
    assert sa.sa().tolist() == v["sa"]
    assert sa.isa().tolist() == v["isa"]
    assert sa.lcp().tolist() == v["lcp"]
    info = sa.info()
    assert info["backend"] == "hip32" and info["coordinate_width"] == 32
    assert info["stored_coordinate_width"] == 32 and info["acceleration"] == 3


def test_sampled_build_matches_fixture(hip, legacy_fixtures):
    # K = 2 (legacy-sa-v1.3): sampled SA, dense ISA, interval-min LCP.
    ref = api.GenomeReference.from_records([("chr1", b"ACGTNACGT"), ("chr2", b"TTTACGTAAA"),
                                            ("chr3", b"AAA")])
    sa = api.SuffixArray.build(ref, api.BuildOptions(sampling_rate=2))
    v = legacy_fixtures["legacy-sa-v1.3"]["vectors"]
    assert sa.sa().tolist() == v["sa"]
    assert sa.isa().tolist() == v["isa"]
    assert sa.lcp().tolist() == v["lcp"]


@pytest.mark.parametrize("k", [2, 3, 8])
def test_sampled_build_matches_generalized_kasai(hip, k):
    text = synth.repeat_text(12, 20000, cover=0.5, n_runs=5, contigs=3)
    ref = api.GenomeReference.from_encoded(text, synth.contig_lengths(text))
    sa = api.SuffixArray.build(ref, api.BuildOptions(sampling_rate=k))
    full = oracle.sa_naive(text)
    sampled = full[full % k == 0]
    assert np.array_equal(sa.sa(), sampled)
    isa = oracle.isa(sampled, k)
    assert np.array_equal(sa.isa(), isa)
    assert np.array_equal(sa.lcp(), oracle.lcp_kasai(text, sampled, isa, k))


def test_count_matches_naive_positions(hip):
    rng = np.random.default_rng(5)
    text = synth.repeat_text(21, 30000, cover=0.5, n_runs=4, contigs=3)
    ref = api.GenomeReference.from_encoded(text, synth.contig_lengths(text))
    sa = api.SuffixArray.build(ref)
    letters = "ACGT"
    raw = "".join("ACGTN\x01\x00"[c - 2] if c >= 2 else "|" for c in text[:-1])
    for _ in range(60):
        if rng.random() < 0.5:
            start = int(rng.integers(0, text.size - 40))
            pattern = "".join(letters[c - 2] if 2 <= c <= 5 else "A" for c in text[start:start + int(rng.integers(1, 25))])
        else:
            pattern = "".join(letters[i] for i in rng.integers(0, 4, size=int(rng.integers(1, 12))))
        want = sum(1 for i in range(len(raw) - len(pattern) + 1) if raw.startswith(pattern, i))
        assert sa.count(pattern) == want, pattern
    with pytest.raises(SufkitError) as e:
        sa.count("ACGN")
    assert e.value.code == "invalid_input"


def test_stage_statistics_reported(hip):
    text = synth.repeat_text(2, 50000)
    ref = api.GenomeReference.from_encoded(text, synth.contig_lengths(text))
    sa = api.SuffixArray.build(ref)
    assert sa.stats["device_seconds"] > 0 and sa.stats["total_seconds"] >= sa.stats["device_seconds"]
    assert np.array_equal(sa.sa(), oracle.fast_build(text)[0])


def test_unsupported_profiles(hip):
    ref = api.GenomeReference.from_records([("a", b"ACGTACGT")])
    with pytest.raises(SufkitError) as e:
        api.SuffixArray.build(ref, api.BuildOptions(acceleration=api.ACCEL_FULL))
    assert e.value.code == "unsupported_backend"
    # Low-memory needs the complete SA (suffix_array.cpp:4143-4146)
    with pytest.raises(SufkitError) as e:
        api.SuffixArray.build(ref, api.BuildOptions(resource_profile=1, sampling_rate=2))
    assert e.value.code == "invalid_input"


@pytest.mark.parametrize("construction", [32, 64])
def test_low_memory_byte_coded_lcp_decodes_to_the_raw_lcp(tmp_path, hip, construction):
    """Low-memory build of a repeat-heavy reference (thousands of LCP values
    of 255 and more): no ISA, byte-coded LCP whose decoded rows equal the
    Fast build's raw LCP, through Save and Load too."""
    rng = np.random.default_rng(construction)
    letters = np.frombuffer(b"ACGT", np.uint8)
    unit = letters[rng.integers(0, 4, 3000)].tobytes()
    parts = []
    for i in range(40):
        parts.append(letters[rng.integers(0, 4, int(rng.integers(100, 5000)))].tobytes())
        parts.append(unit[int(rng.integers(0, 1000)):])
        if i % 7 == 0:
            parts.append(b"A" * int(rng.integers(300, 2000)) + b"N" * int(rng.integers(300, 900)))
    ref = api.GenomeReference.from_records([("c1", b"".join(parts[:40])), ("c2", b"".join(parts[40:]))])
    fast = api.SuffixArray.build(ref, api.BuildOptions(backend=api.BACKEND_HIP, coordinate_width=construction))
    low = api.SuffixArray.build(ref, api.BuildOptions(backend=api.BACKEND_HIP, coordinate_width=construction,
                                                      resource_profile=1))
    raw = fast.lcp()
    assert int((raw >= 255).sum()) > 1000
    info = low.info()
    assert info["lcp_encoding"] == 2 and low.isa() is None
    assert np.array_equal(low.sa(), fast.sa())
    assert np.array_equal(low.lcp(), raw)
    path = tmp_path / "low.sufidx"
    low.save(path)
    back = api.SuffixArray.load(path)
    assert np.array_equal(back.lcp(), raw)
    assert back.info()["lcp_bytes"] == info["lcp_bytes"]
    assert info["lcp_bytes"] < raw.size * 2


def test_save_refuses_to_clobber(hip, tmp_path):
    ref = api.GenomeReference.from_records([("a", b"ACGTACGTTT")])
    sa = api.SuffixArray.build(ref)
    path = tmp_path / "x.sufidx"
    sa.save(path)
    with pytest.raises(SufkitError) as e:
        sa.save(path)
    assert e.value.code == "io_error"
    sa.save(path, overwrite=True)
    assert not list(tmp_path.glob("*.partial.*"))

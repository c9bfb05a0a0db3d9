"""Host library on the CPU: reference normalization, width policies, and
loading .sufidx files written by the reference itself (legacy 1.0-1.3
fixtures of tests/data and 1.4 files from the reference's own Save).

Patterns follow test_sufkit.cpp:88-190 (normalization and offsets),
test_caps_backend.cpp:169-236 (selection policy), test_sufkit.cpp:913-941
(legacy fixtures) and test_sa_codec.cpp (corruption)."""
import gzip
import json
import shutil
import struct
from pathlib import Path

import numpy as np
import pytest

import oracle
from sufkit_amd import SufkitError, api
from tests.golden.make_golden import golden_references

GOLDEN = Path(__file__).resolve().parent / "golden"
FASTA = ">chr1 first\nACGTNACGT\n>chr2 second\nTTTACGTAAA\n>chr3 third\nAAA\n"


def test_from_records_normalization_and_fingerprint(legacy_fixtures):
    ref = api.GenomeReference.from_records([("chr1", b"ACGTNACGT"), ("chr2", b"TTTACGTAAA"),
                                            ("chr3", b"AAA")])
    fx = legacy_fixtures["legacy-sa-v1.1"]
    s = ref.summary()
    assert s["sequence_count"] == 3 and s["total_bases"] == fx["total_bases"] == 22
    assert s["text_symbols"] == fx["text_symbols"] == 26
    assert s["ambiguous_bases"] == 1
    assert f"0x{s['fingerprint']:016x}" == fx["fingerprint"]
    assert ref.text().tolist() == fx["vectors"]["text"]
    # global offsets: +len+1 per contig (chr2 starts at 10)
    assert ref.contigs()[:, 0].tolist() == [0, 10, 21]


def test_from_fasta_plain_and_gzip(tmp_path, legacy_fixtures):
    plain = tmp_path / "ref.fa"
    plain.write_text(FASTA)
    gz = tmp_path / "ref.fa.gz"
    with gzip.open(gz, "wt") as f:
        f.write(FASTA.replace("ACGTNACGT\n", "ACGTN\nACGT\n"))  # wrapped lines
    for path in (plain, gz):
        ref = api.GenomeReference.from_fasta(path)
        assert ref.text().tolist() == legacy_fixtures["legacy-sa-v1.1"]["vectors"]["text"]


def test_lowercase_and_iupac_normalize_like_oracle():
    seqs = [b"acgtRYKMswbdhvN-.", b"GATTACA"]
    ref = api.GenomeReference.from_records([("a", seqs[0]), ("b", seqs[1])])
    text, amb = oracle.encode_records(seqs)
    assert np.array_equal(ref.text(), text)
    assert ref.summary()["ambiguous_bases"] == amb
    assert ref.summary()["fingerprint"] == oracle.fingerprint(text[:-1])


@pytest.mark.parametrize("records,needle", [
    ([], "one record or more"),
    ([("a", b"ACGT"), ("a", b"GG")], "share the name"),
    ([("", b"ACGT")], "has no name"),
    ([("a", b"")], "holds no bases"),
])
def test_reference_validation_errors(records, needle):
    with pytest.raises(SufkitError) as e:
        api.GenomeReference.from_records(records)
    assert e.value.code == "invalid_input" and needle in str(e.value)


def test_missing_fasta_is_io_error(tmp_path):
    with pytest.raises(SufkitError) as e:
        api.GenomeReference.from_fasta(tmp_path / "absent.fa")
    assert e.value.code == "io_error"


HIP, DIV, CAPS, AUTO = api.BACKEND_HIP, api.BACKEND_DIVSUFSORT, api.BACKEND_CAPS, api.BACKEND_AUTO


@pytest.mark.parametrize("args,want", [
    # auto construction width: 32 up to UINT32_MAX symbols (like CaPS32)
    ((AUTO, 0, 0, 0, 26), dict(backend=HIP, construction_width=32, stored_width=32, isa_width=32, lcp_width=32)),
    ((HIP, 0, 0, 0, 2**32 - 1), dict(construction_width=32, stored_width=32, lcp_width=32)),
    ((HIP, 0, 0, 0, 2**32), dict(construction_width=64, stored_width=32, lcp_width=32)),
    ((HIP, 0, 0, 0, 2**32 + 1), dict(construction_width=64, stored_width=64, lcp_width=64)),
    ((HIP, 64, 0, 0, 1000), dict(construction_width=64, stored_width=32)),
    ((HIP, 0, 40, 0, 1000), dict(stored_width=40, isa_width=40)),
    ((HIP, 0, 48, 0, 1000), dict(stored_width=48)),
    ((HIP, 0, 0, 1, 2**32 + 1), dict(stored_width=40)),   # low-memory picks the narrowest
    ((DIV, 0, 0, 0, 2**31 - 1), dict(construction_width=32)),  # divsufsort32 is signed
    ((DIV, 0, 0, 0, 2**31), dict(construction_width=64)),
])
def test_build_policy(args, want):
    got = api.resolve_build_policy(*args)
    for key, value in want.items():
        assert got[key] == value, (key, got)


@pytest.mark.parametrize("args", [
    (HIP, 32, 0, 0, 2**32),          # too large for 32-bit construction
    (HIP, 0, 32, 0, 2**32 + 1),      # explicit storage width too narrow
])
def test_build_policy_rejects(args):
    with pytest.raises(SufkitError) as e:
        api.resolve_build_policy(*args)
    assert e.value.code == "invalid_input"


# ------------------------------------------------------------- loading ---

def write_fixture(tmp_path, legacy_fixtures, name):
    fx = legacy_fixtures[name]
    header = bytearray.fromhex(fx["header_zeroed_crc_hex"])
    struct.pack_into("<I", header, 72, fx["header_crc"])
    data = bytearray(header)
    for section in fx["sections"]:
        data += bytes.fromhex(section["hex"])
    path = tmp_path / f"{name}.sufidx"
    path.write_bytes(bytes(data))
    return path


@pytest.mark.parametrize("name", ["legacy-sa-v1.0", "legacy-sa-v1.1", "legacy-sa-v1.2", "legacy-sa-v1.3"])
def test_load_reference_legacy_fixtures(tmp_path, legacy_fixtures, name):
    path = write_fixture(tmp_path, legacy_fixtures, name)
    sa = api.SuffixArray.load(path)
    v = legacy_fixtures[name]["vectors"]
    assert sa.sa().tolist() == v["sa"]
    if "isa" in v:
        assert sa.isa().tolist() == v["isa"]
    if "lcp" in v:
        assert sa.lcp().tolist() == v["lcp"]
    info = sa.info()
    assert info["backend"] == "divsufsort32" and info["coordinate_width"] == 32
    assert info["format_version"] == legacy_fixtures[name]["format"]
    assert info["sa_sampling_rate"] == v.get("sampling_rate", 1)


def golden_files():
    index = json.loads((GOLDEN / "sufidx_index.json").read_text())
    return sorted(index.items())


@pytest.mark.parametrize("fname,meta", golden_files())
def test_load_reference_written_v14_files(fname, meta):
    """Files from the reference's own SuffixArray::Save: our loader accepts
    them and their payloads equal the oracle's SA/ISA/LCP."""
    sa = api.SuffixArray.load(GOLDEN / "sufidx" / fname)
    records = golden_references()[meta["reference"]]
    text, _ = oracle.encode_records([s for _, s in records])
    osa, oisa, olcp = oracle.fast_build(text, "naive")
    assert np.array_equal(sa.sa(), osa)
    info = sa.info()
    assert info["format_version"] == "1.4"
    assert info["coordinate_width"] == (meta["coordinate_width"] or 32)
    assert info["stored_coordinate_width"] == (meta["storage_width"] or 32)
    if meta.get("resource_profile", 0) == 1:
        # Low-memory: no ISA, byte-coded LCP decoded through its anchors
        assert sa.isa() is None
        assert info["lcp_encoding"] == 2
        assert np.array_equal(sa.lcp(), olcp)
        return
    if meta["acceleration"] == 3:
        assert np.array_equal(sa.isa(), oisa)
    if meta["acceleration"] in (1, 3):
        assert np.array_equal(sa.lcp(), olcp)
    else:
        assert sa.lcp() is None


def test_corrupted_files_are_rejected(tmp_path):
    src = GOLDEN / "sufidx" / "mixed_cauto_sauto_a3.sufidx"
    data = bytearray(src.read_bytes())
    cases = {}
    flipped = bytearray(data)
    flipped[len(data) // 2] ^= 0x01            # payload byte: section CRC
    cases["payload"] = flipped
    header = bytearray(data)
    header[40] ^= 0x01                         # header field: header CRC
    cases["header"] = header
    cases["truncated"] = data[: len(data) - 17]
    for label, blob in cases.items():
        path = tmp_path / f"{label}.sufidx"
        path.write_bytes(bytes(blob))
        with pytest.raises(SufkitError) as e:
            api.SuffixArray.load(path)
        assert e.value.code == "corrupt_index", label


def test_unknown_backend_id_rejected(tmp_path):
    data = bytearray((GOLDEN / "sufidx" / "fixture_cauto_sauto_a3.sufidx").read_bytes())
    header_size = struct.unpack_from("<I", data, 28)[0]
    data[14] = 42                                # unknown stored backend
    data[72:76] = b"\0\0\0\0"
    struct.pack_into("<I", data, 72, oracle.crc32(bytes(data[:header_size])))
    path = tmp_path / "unknown.sufidx"
    path.write_bytes(bytes(data))
    with pytest.raises(SufkitError) as e:
        api.SuffixArray.load(path)
    assert e.value.code == "unsupported_backend"


def test_future_minor_version_rejected(tmp_path):
    data = bytearray((GOLDEN / "sufidx" / "fixture_cauto_sauto_a3.sufidx").read_bytes())
    struct.pack_into("<H", data, 10, 9)
    path = tmp_path / "future.sufidx"
    path.write_bytes(bytes(data))
    with pytest.raises(SufkitError) as e:
        api.SuffixArray.load(path)
    assert e.value.code == "version_mismatch"


def test_build_without_device_is_unsupported_backend():
    if api.hip_backend_available():
        pytest.skip("a HIP device is visible")
    ref = api.GenomeReference.from_records([("a", b"ACGT")])
    with pytest.raises(SufkitError) as e:
        api.SuffixArray.build(ref)
    assert e.value.code == "unsupported_backend"


def test_cpu_engines_are_not_part_of_this_build():
    ref = api.GenomeReference.from_records([("a", b"ACGT")])
    for backend in (api.BACKEND_DIVSUFSORT, api.BACKEND_CAPS):
        with pytest.raises(SufkitError) as e:
            api.SuffixArray.build(ref, api.BuildOptions(backend=backend))
        assert e.value.code == "unsupported_backend"


def test_build_option_validation():
    ref = api.GenomeReference.from_records([("a", b"ACGT")])
    for opts in (api.BuildOptions(threads=0), api.BuildOptions(sampling_rate=0)):
        with pytest.raises(SufkitError) as e:
            api.SuffixArray.build(ref, opts)
        assert e.value.code == "invalid_input"

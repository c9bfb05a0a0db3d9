"""Write the golden vectors under tests/golden/ (run in the build container).

  legacy_fixtures.json
      Decoded from the reference's own data files tests/data/legacy-sa-v1.*.
      fixture (.sufidx 1.0-1.3 written by sufkit commit fd1abbd from
      tests/data/reference.fa): header fields, every section's bytes and
      CRC32, and the text/SA/ISA/LCP/CHILD vectors they hold.  Only the data
      is read; no reference code is run.
  reference_digests.json
      SA/ISA/LCP digests of the reference's own divsufsort backend
      (oracle/_ref/libsufkit_ref.so, compiled from the reference sources) on
      deterministic synthetic texts, for parity at sizes the GPU tests run
      and the oracle cannot finish quickly.

Usage: python tests/golden/make_golden.py [--reference /root/reference]
"""
from __future__ import annotations

import argparse
import json
import struct
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

SECTION_NAMES = {1: "metadata", 2: "text", 3: "sa", 5: "isa", 6: "lcp", 7: "child",
                 8: "learned", 9: "sampling"}


def decode_fixture(data: bytes) -> dict:
    magic = data[:8]
    assert magic == b"SUFKIDX\x00"
    (major, minor, endian, kind, backend, width, normalization, v0, v1, v2, s0, s1, s2,
     profile, section_count, header_size, seq_count, total_bases, text_symbols, ambiguous,
     fingerprint, header_crc, _reserved) = struct.unpack_from("<HHBBBBBBBBBBBBIIQQQQQII", data, 8)
    header = bytearray(data[:header_size])
    header[72:76] = b"\0\0\0\0"
    out = {
        "format": f"{major}.{minor}", "kind": kind, "backend": backend,
        "construction_width": width, "normalization": normalization,
        "library_version": [v0, v1, v2], "profile": profile,
        "sequence_count": seq_count, "total_bases": total_bases,
        "text_symbols": text_symbols, "ambiguous_bases": ambiguous,
        "fingerprint": f"0x{fingerprint:016x}", "header_crc": header_crc,
        "header_zeroed_crc_hex": bytes(header).hex(), "sections": [],
    }
    vectors = {}
    for i in range(section_count):
        stype, flags, offset, size, crc, _ = struct.unpack_from("<IIQQII", data, 80 + 32 * i)
        payload = data[offset:offset + size]
        out["sections"].append({"type": stype, "name": SECTION_NAMES.get(stype, str(stype)),
                                "flags": flags, "crc": crc, "hex": payload.hex()})
        if stype == 2:
            vectors["text"] = list(payload)
        elif stype == 3:
            count = struct.unpack_from("<Q", payload, 0)[0]
            w = width // 8
            vectors["sa"] = [int.from_bytes(payload[8 + w * k:8 + w * (k + 1)], "little")
                             for k in range(count)]
        elif stype in (5, 6, 7):
            count = struct.unpack_from("<Q", payload, 0)[0]
            w = payload[8] // 8
            vectors[SECTION_NAMES[stype]] = [
                int.from_bytes(payload[9 + w * k:9 + w * (k + 1)], "little") for k in range(count)]
        elif stype == 9:
            vectors["sampling_rate"] = struct.unpack_from("<I", payload, 0)[0]
    out["vectors"] = vectors
    return out


def legacy(reference: Path) -> dict:
    result = {}
    for v in range(4):
        path = reference / "tests" / "data" / f"legacy-sa-v1.{v}.fixture"
        result[f"legacy-sa-v1.{v}"] = decode_fixture(path.read_bytes())
    fa = (reference / "tests" / "data" / "reference.fa").read_text()
    result["reference.fa"] = fa
    return result


def reference_digests() -> dict:
    import oracle
    from tests import synth  # deterministic generators shared with the tests

    out = {}
    for name, text in synth.golden_cases():
        for method in ("ref32", "ref64"):
            sa, isa, lcp = oracle.fast_build(text, method)
            out.setdefault(name, {})[method] = {
                "n": int(text.size), "fingerprint": f"0x{oracle.fingerprint(text[:-1]):016x}",
                "sa": f"0x{oracle.digest(sa):016x}", "isa": f"0x{oracle.digest(isa):016x}",
                "lcp": f"0x{oracle.digest(lcp):016x}",
                "sa_head": [int(x) for x in sa[:16]], "lcp_max": int(lcp.max()),
            }
    return out


def golden_references() -> dict[str, list[tuple[str, bytes]]]:
    """Small references for the .sufidx golden files (records as FASTA
    text would give them)."""
    import numpy as np

    rng = np.random.default_rng(20260822)
    alphabet = np.frombuffer(b"ACGT", np.uint8)
    def rand(k):
        return alphabet[rng.integers(0, 4, size=k)].tobytes()
    mixed = []
    for c in range(5):
        body = bytearray(rand(int(rng.integers(500, 4000))))
        pos = int(rng.integers(0, len(body) - 200))
        body[pos:pos + 150] = b"N" * 150                     # N run
        body[10:90] = b"A" * 80                              # poly-A
        body.extend(b"acgtRYKM" + rand(300))                 # lowercase + IUPAC -> N
        mixed.append((f"chr{c + 1}", bytes(body)))
    # Long exact repeats, a long poly-A and a long N run: LCP values of 255
    # and more, so a byte-coded LCP needs overflow anchors (runs that continue
    # and runs that restart).
    rep_rng = np.random.default_rng(20261203)
    def rep_rand(k):
        return alphabet[rep_rng.integers(0, 4, size=k)].tobytes()
    unit = rep_rand(1500)
    repeats = [
        ("rep1", rep_rand(400) + unit + rep_rand(200) + unit + b"A" * 700 + rep_rand(300) + unit[:900]),
        ("rep2", rep_rand(250) + b"N" * 600 + unit[300:] + rep_rand(500) + b"AC" * 400),
    ]
    return {
        "fixture": [("chr1", b"ACGTNACGT"), ("chr2", b"TTTACGTAAA"), ("chr3", b"AAA")],
        "mixed": mixed,
        "repeats": repeats,
    }


# (coordinate width, storage width, acceleration) variants written per reference
SUFIDX_VARIANTS = [
    (0, 0, 3), (64, 0, 3), (0, 40, 3), (0, 48, 3), (0, 64, 3), (0, 0, 1), (0, 0, 0), (64, 48, 1),
]


# Low-memory profile variants (byte-coded LCP, no ISA; the profile sets the
# acceleration to LCP whatever is asked).
LOWMEM_VARIANTS = [(0, 0, 3), (64, 0, 3), (0, 40, 3), (0, 64, 0)]


def sufidx_golden() -> dict:
    import oracle

    out_dir = HERE / "sufidx"
    out_dir.mkdir(exist_ok=True)
    index = {}
    for name, records in golden_references().items():
        for cw, sw, accel in SUFIDX_VARIANTS:
            fname = f"{name}_c{cw or 'auto'}_s{sw or 'auto'}_a{accel}.sufidx"
            oracle.ref_write_sufidx(records, str(out_dir / fname), cw, sw, accel)
            index[fname] = {"reference": name, "coordinate_width": cw, "storage_width": sw,
                            "acceleration": accel, "resource_profile": 0}
        for cw, sw, accel in LOWMEM_VARIANTS:
            fname = f"{name}_c{cw or 'auto'}_s{sw or 'auto'}_a{accel}_lowmem.sufidx"
            oracle.ref_write_sufidx(records, str(out_dir / fname), cw, sw, accel, resource_profile=1)
            index[fname] = {"reference": name, "coordinate_width": cw, "storage_width": sw,
                            "acceleration": accel, "resource_profile": 1}
    return index


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    args = ap.parse_args()
    ref = Path(args.reference)
    (HERE / "legacy_fixtures.json").write_text(json.dumps(legacy(ref), indent=1) + "\n")
    import oracle
    oracle.build(str(ref))
    (HERE / "reference_digests.json").write_text(json.dumps(reference_digests(), indent=1) + "\n")
    if oracle.ref_writer_available():
        (HERE / "sufidx_index.json").write_text(json.dumps(sufidx_golden(), indent=1) + "\n")


if __name__ == "__main__":
    main()

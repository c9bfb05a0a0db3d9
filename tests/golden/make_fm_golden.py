"""Write tests/golden/fm_golden.json: query results of the reference's own
FmIndex (SDSL csa_wt built on the host, oracle/_ref/libsufkit_refwriter.so
compiled from the reference sources) on small deterministic references.

For each reference: its records, a pattern list, and per pattern the
EqualRange, Count for the three strand modes, and Locate (three strand modes;
every hit when there are few, max_hits 0 and 3 always); CountBatch over the
whole list for each strand mode and two batch widths; the error codes
(ErrorCode + 1) of invalid calls.  Only inputs and outputs are stored.

Usage (build container, reference present):
    python tests/golden/make_fm_golden.py
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402

PALINDROMES = ["AT", "TA", "GC", "CG", "ACGT", "AATT", "GAATTC", "TTAA"]
LENGTHS = [1, 2, 3, 5, 8, 10, 11, 12, 16, 20, 31, 32, 33, 40]


def references() -> list[list[tuple[str, str]]]:
    rng = np.random.default_rng(20260822)
    small = [("chr1", "ACGTACGTTTGACCA"), ("chr2", "NNACGTAC"), ("chr3", "GGTACGTACGT")]

    def random_seq(size: int) -> str:
        return "".join("ACGT"[c] for c in rng.integers(0, 4, size))

    unit = random_seq(300)
    contigs = []
    for c in range(5):
        parts = []
        for _ in range(6):
            parts.append(random_seq(int(rng.integers(200, 900))))
            copy = list(unit)
            for j in rng.integers(0, len(copy), 6):
                copy[j] = "ACGT"[int(rng.integers(0, 4))]
            parts.append("".join(copy))
            if rng.random() < 0.5:
                parts.append("N" * int(rng.integers(5, 60)))
        parts.append("AT" * 40 + "A" * 50)
        contigs.append((f"ctg{c}", "".join(parts)))
    return [small, contigs]


def patterns_for(records: list[tuple[str, str]], seed: int) -> list[str]:
    rng = np.random.default_rng(seed)
    text = "".join(seq for _, seq in records)
    out = list(PALINDROMES)
    for length in LENGTHS:
        for _ in range(3):
            at = int(rng.integers(0, max(1, len(text) - length)))
            window = text[at:at + length].replace("N", "A")
            if len(window) == length:
                out.append(window)
        out.append("".join("ACGT"[c] for c in rng.integers(0, 4, length)))
    out.append("acgt")  # lower case is accepted
    return out


def main() -> int:
    if not oracle.ref_fm_available():
        print("oracle/_ref/libsufkit_refwriter.so is not built", file=sys.stderr)
        return 1
    cases = []
    for i, records in enumerate(references()):
        raw = [(name, seq.encode()) for name, seq in records]
        fm = oracle.RefFmIndex(raw)
        assert fm.build_status == 0
        patterns = patterns_for(records, 100 + i)
        per_pattern = []
        for p in patterns:
            entry = {"pattern": p, "equal_range": list(fm.equal_range(p)),
                     "count": [fm.count(p, s) for s in range(3)], "locate": {}}
            for s in range(3):
                for limit in (None, 0, 3):
                    if limit is None and fm.count(p, s) > 64:
                        continue
                    entry["locate"][f"{s}:{limit}"] = fm.locate(p, s, limit)
            per_pattern.append(entry)
        batches = {f"{s}:{w}": fm.count_batch(patterns, s, w) for s in range(3) for w in (0, 7)}
        errors = {
            "equal_range_empty": fm.equal_range("")[1],
            "equal_range_non_acgt": fm.equal_range("ACGN")[1],
            "count_bad_strand": fm.count("ACGT", 3)[1],
            "count_batch_width": fm.count_batch(["ACGT"], 0, 300)[1],
            "count_batch_bad_pattern": fm.count_batch(["ACGT", "AXG"], 0, 0)[1],
            "locate_bad_strand": fm.locate("ACGT", 3)[1],
        }
        sada = oracle.RefFmIndex(raw, backend=3)
        errors["build_sada"] = sada.build_status
        cases.append({"records": records, "patterns": per_pattern, "count_batch": batches, "errors": errors})
    out = HERE / "fm_golden.json"
    out.write_text(json.dumps({"source": "reference FmIndex (SDSL csa_wt_huff) via oracle/_ref", "cases": cases},
                              separators=(",", ":")))
    print(f"wrote {out} ({out.stat().st_size} bytes)")
    return 0


if __name__ == "__main__":
    sys.exit(main())

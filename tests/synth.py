"""Deterministic test texts (logical texts: encoded bytes + sentinel 0).

Symbols follow the reference encoding (reference_data.hpp:20-27):
0 sentinel, 1 separator, 2 A, 3 C, 4 G, 5 T, 6 N.  Every generator returns
a uint8 array whose last element is the only 0.
"""
from __future__ import annotations

import numpy as np

A, C, G, T, N, SEP = 2, 3, 4, 5, 6, 1


def finish(contigs: list[np.ndarray]) -> np.ndarray:
    """Concatenate contigs with a separator after each, then the sentinel."""
    parts = []
    for c in contigs:
        parts.append(np.asarray(c, dtype=np.uint8))
        parts.append(np.array([SEP], dtype=np.uint8))
    parts.append(np.array([0], dtype=np.uint8))
    return np.concatenate(parts)


def contig_lengths(text: np.ndarray) -> np.ndarray:
    """Contig lengths of a logical text (bases between separators)."""
    seps = np.flatnonzero(text == SEP)
    starts = np.r_[0, seps[:-1] + 1]
    return (seps - starts).astype(np.uint64)


def random_bases(rng: np.random.Generator, length: int) -> np.ndarray:
    return rng.integers(A, T + 1, size=length, dtype=np.uint8)


def random_text(seed: int, length: int) -> np.ndarray:
    """One contig of i.i.d. uniform A/C/G/T."""
    rng = np.random.default_rng(seed)
    return finish([random_bases(rng, length)])


def random_symbols(seed: int, length: int) -> np.ndarray:
    """Uniform over A/C/G/T/N/separator (like the 6-symbol texts of
    test_caps_backend.cpp:43-94), sentinel appended."""
    rng = np.random.default_rng(seed)
    body = rng.integers(1, 7, size=length, dtype=np.uint8)
    return np.concatenate([body, np.array([0], np.uint8)])


def run_text(seed: int, pieces: list[tuple[str, int]]) -> np.ndarray:
    """Concatenate pieces: ('rand', L), ('A', L) poly runs, ('N', L),
    ('tandem2', L), ('tandem7', L), ('sep', 1)."""
    rng = np.random.default_rng(seed)
    contigs: list[np.ndarray] = []
    cur: list[np.ndarray] = []
    for kind, length in pieces:
        if kind == "rand":
            cur.append(random_bases(rng, length))
        elif kind in ("A", "C", "G", "T", "N"):
            cur.append(np.full(length, {"A": A, "C": C, "G": G, "T": T, "N": N}[kind], np.uint8))
        elif kind.startswith("tandem"):
            period = int(kind[6:])
            unit = random_bases(rng, period)
            cur.append(np.resize(unit, length))
        elif kind == "sep":
            contigs.append(np.concatenate(cur) if cur else np.zeros(0, np.uint8))
            cur = []
        else:
            raise ValueError(kind)
    if cur:
        contigs.append(np.concatenate(cur))
    contigs = [c for c in contigs if c.size]
    return finish(contigs)


def repeat_text(seed: int, length: int, families: int = 8, family_len: tuple[int, int] = (300, 3000),
                cover: float = 0.4, divergence: tuple[float, float] = (0.0, 0.05),
                n_runs: int = 4, contigs: int = 3) -> np.ndarray:
    """Random genome with copied repeat families (with substitutions), N runs
    and several contigs."""
    rng = np.random.default_rng(seed)
    body = random_bases(rng, length)
    fams = [random_bases(rng, int(rng.integers(*family_len))) for _ in range(families)]
    covered = 0
    while covered < cover * length:
        fam = fams[int(rng.integers(len(fams)))]
        if fam.size >= length:
            break
        copy = fam.copy()
        div = rng.uniform(*divergence)
        mask = rng.random(copy.size) < div
        copy[mask] = rng.integers(A, T + 1, size=int(mask.sum()), dtype=np.uint8)
        pos = int(rng.integers(0, length - copy.size))
        body[pos:pos + copy.size] = copy
        covered += copy.size
    for _ in range(n_runs):
        run = int(rng.integers(10, 400))
        pos = int(rng.integers(0, max(1, length - run)))
        body[pos:pos + run] = N
    cuts = sorted(int(x) for x in rng.integers(1, length, size=contigs - 1)) if contigs > 1 else []
    pieces = np.split(body, cuts)
    return finish([p for p in pieces if p.size])


def fixture_text() -> np.ndarray:
    """FromRecords of tests/data/reference.fa (chr1 ACGTNACGT, chr2
    TTTACGTAAA, chr3 AAA) + sentinel: the 26-symbol text of the legacy
    fixtures (test_sufkit.cpp:220-243)."""
    table = {"A": A, "C": C, "G": G, "T": T}
    contigs = [np.array([table.get(ch, N) for ch in s], np.uint8)
               for s in ("ACGTNACGT", "TTTACGTAAA", "AAA")]
    return finish(contigs)


def golden_cases() -> list[tuple[str, np.ndarray]]:
    """Medium texts whose reference digests are committed in
    tests/golden/reference_digests.json."""
    return [
        ("random_1M", random_text(20260822, 1_000_000)),
        ("repeats_600k", repeat_text(7, 600_000, cover=0.5, divergence=(0.0, 0.03))),
        ("runs_300k", run_text(11, [("rand", 50_000), ("A", 40_000), ("rand", 20_000),
                                    ("tandem2", 30_000), ("sep", 1), ("tandem7", 35_000),
                                    ("N", 25_000), ("rand", 50_000), ("sep", 1),
                                    ("N", 5_000), ("rand", 40_000), ("C", 5_000)])),
        ("symbols_200k", random_symbols(3, 200_000)),
    ]

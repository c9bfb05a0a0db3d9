"""Batched build: many references in one device build, each equal to its own
single build and to the oracle, bit for bit."""
import numpy as np
import pytest

import oracle
import sufkit_amd
from tests import synth

pytestmark = pytest.mark.gpu


def batch_texts(seed, count):
    rng = np.random.default_rng(seed)
    texts = []
    for r in range(count):
        kind = r % 5
        size = int(rng.integers(1, 4000))
        if kind == 0:
            texts.append(synth.random_text(seed * 100 + r, size))
        elif kind == 1:
            texts.append(synth.random_symbols(seed * 100 + r, size))
        elif kind == 2:
            texts.append(synth.run_text(r, [("rand", 30), ("A", size), ("sep", 1), ("tandem3", 200)]))
        elif kind == 3:
            texts.append(synth.repeat_text(r, max(size, 400), cover=0.6, n_runs=2, contigs=2))
        else:
            texts.append(synth.random_text(r, 1 + r % 3))  # tiny references
    return texts


@pytest.mark.parametrize("bits", [32, 64])
def test_batch_equals_oracle_per_reference(hip, bits):
    texts = batch_texts(3, 40)
    parts, stats = hip.hip_build_batch(texts, construction_bits=bits, out_bits=32)
    assert len(parts) == len(texts)
    for text, (sa, isa, lcp) in zip(texts, parts):
        osa, oisa, olcp = oracle.fast_build(text)
        assert np.array_equal(sa.astype(np.uint64), osa)
        assert np.array_equal(isa.astype(np.uint64), oisa)
        assert np.array_equal(lcp.astype(np.uint64), olcp)


def test_batch_equals_single_builds(hip):
    texts = [sufkit_amd.synthetic_text(4, 20260822, ref, 0.05)[0] for ref in range(12)]
    parts, _ = hip.hip_build_batch(texts)
    for text, (sa, isa, lcp) in zip(texts, parts):
        ssa, sisa, slcp, _ = hip.hip_build(text)
        assert np.array_equal(sa, ssa) and np.array_equal(isa, sisa) and np.array_equal(lcp, slcp)


def test_single_reference_batch(hip):
    text = synth.repeat_text(1, 30000)
    parts, _ = hip.hip_build_batch([text])
    sa, isa, lcp, _ = hip.hip_build(text)
    assert np.array_equal(parts[0][0], sa) and np.array_equal(parts[0][2], lcp)


def check_against_oracle(texts, parts):
    for text, (sa, isa, lcp) in zip(texts, parts):
        osa, oisa, olcp = oracle.fast_build(text)
        assert np.array_equal(sa.astype(np.uint64), osa)
        assert np.array_equal(isa.astype(np.uint64), oisa)
        assert np.array_equal(lcp.astype(np.uint64), olcp)


def test_many_tiny_references(hip):
    # Hundreds of references inside one radix tile and one refine tile.
    rng = np.random.default_rng(11)
    texts = [synth.random_text(9000 + r, int(rng.integers(1, 30))) for r in range(5000)]
    parts, _ = hip.hip_build_batch(texts)
    check_against_oracle(texts, parts)


def test_more_references_than_the_placement_tables(hip):
    # Beyond 65536 references the first step sorts (reference, key) pairs
    # instead of sorting in place per reference.
    rng = np.random.default_rng(12)
    texts = [synth.random_text(20000 + r, int(rng.integers(1, 7))) for r in range(70000)]
    parts, _ = hip.hip_build_batch(texts)
    check_against_oracle(texts, parts)


def test_batch_rejects_bad_layouts(hip):
    good = [synth.random_text(1, 50), synth.random_text(2, 60)]
    bad_sentinel = [good[0].copy(), good[1].copy()]
    bad_sentinel[0][-1] = 2          # reference 0 lacks its sentinel
    bad_middle = [good[0].copy(), good[1].copy()]
    bad_middle[1][5] = 0             # a second 0 inside reference 1
    for texts in (bad_sentinel, bad_middle):
        with pytest.raises(sufkit_amd.SufkitError) as e:
            hip.hip_build_batch(texts)
        assert e.value.code == "invalid_input"
    builder = hip.DeviceBuilder(0)
    d = hip.DeviceArray(256)
    with pytest.raises(sufkit_amd.SufkitError) as e:
        builder.build_batch(d.ptr, 111, np.array([0, 50, 100], np.uint64), 32, d.ptr, 32, None, 32, None, 32)
    assert e.value.code == "invalid_input"   # offsets do not end at n
    builder.close()

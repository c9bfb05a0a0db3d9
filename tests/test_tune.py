"""Radix shape autotuner: the shape table (CPU) and exact builds under every
compiled shape and after tuning (GPU)."""
import numpy as np
import pytest

import oracle
import sufkit_amd
from tests import synth

KINDS = range(6)
CLASSES = range(7)
SHAPES = [(256, 16), (512, 8), (512, 16), (1024, 8), (1024, 12), (1024, 16), (512, 24)]


def cells():
    """Every (kind, class, digit width) of the table: the per-reference kinds
    4 and 5 sort with 8-bit digits only."""
    return {(k, c, b) for k in KINDS for c in CLASSES for b in (8, 7) if b == 8 or k < 4}


def test_defaults_cover_every_kind_class_and_width():
    sufkit_amd.tune_reset()
    table = sufkit_amd.tune_get()
    assert len(table) == 70
    assert {(e["kind"], e["size_class"], e["digit_bits"]) for e in table} == cells()
    assert all(e["tuned"] == 0 and e["us_per_pass"] == 0 and (e["block"], e["items"]) in SHAPES for e in table)


def test_set_round_trip_and_reset():
    sufkit_amd.tune_reset()
    sufkit_amd.tune_set([{"kind": 1, "size_class": 3, "block": 512, "items": 16, "chosen": 1},
                         {"kind": 1, "size_class": 3, "block": 1024, "items": 8, "chosen": 1, "digit_bits": 7,
                          "us_per_pass": 2.5}])
    table = {(e["kind"], e["size_class"], e["digit_bits"]): e for e in sufkit_amd.tune_get()}
    assert (table[(1, 3, 8)]["block"], table[(1, 3, 8)]["items"], table[(1, 3, 8)]["tuned"]) == (512, 16, 1)
    assert (table[(1, 3, 7)]["block"], table[(1, 3, 7)]["items"], table[(1, 3, 7)]["tuned"]) == (1024, 8, 1)
    assert table[(1, 3, 7)]["us_per_pass"] == 2.5 and table[(1, 3, 8)]["us_per_pass"] == 0
    assert table[(1, 2, 8)]["tuned"] == 0 and table[(1, 2, 7)]["tuned"] == 0
    sufkit_amd.tune_reset()
    assert all(e["tuned"] == 0 and e["us_per_pass"] == 0 for e in sufkit_amd.tune_get())


def test_digit_width_zero_reads_as_eight():
    """Tables saved before the widths carry no digit_bits (0 in the struct)."""
    sufkit_amd.tune_reset()
    try:
        sufkit_amd.tune_set([{"kind": 2, "size_class": 1, "block": 256, "items": 16, "chosen": 1, "digit_bits": 0}])
        table = {(e["kind"], e["size_class"], e["digit_bits"]): e for e in sufkit_amd.tune_get()}
        assert table[(2, 1, 8)]["tuned"] == 1 and table[(2, 1, 7)]["tuned"] == 0
    finally:
        sufkit_amd.tune_reset()


@pytest.mark.parametrize("entry", [
    {"kind": 0, "size_class": 0, "block": 333, "items": 8, "chosen": 1},    # no such shape
    {"kind": 3, "size_class": 0, "block": 1024, "items": 16, "chosen": 1},  # u64/u64 tile exceeds the LDS
    {"kind": 6, "size_class": 0, "block": 512, "items": 8, "chosen": 1},    # no such kind
    {"kind": 0, "size_class": 7, "block": 512, "items": 8, "chosen": 1},    # no such class
    {"kind": 0, "size_class": 0, "block": 512, "items": 8, "chosen": 1, "digit_bits": 6},  # no such width
    {"kind": 4, "size_class": 0, "block": 512, "items": 8, "chosen": 1, "digit_bits": 7},  # per reference: 8 only
    {"kind": 5, "size_class": 2, "block": 512, "items": 8, "chosen": 1, "digit_bits": 7},
])
def test_set_rejects_shapes_that_are_not_compiled(entry):
    sufkit_amd.tune_reset()
    with pytest.raises(sufkit_amd.SufkitError) as e:
        sufkit_amd.tune_set([entry])
    assert e.value.code == "invalid_input"
    assert all(e["tuned"] == 0 for e in sufkit_amd.tune_get())


def install_everywhere(block, items, digit_bits=8, us_per_pass=0.0):
    """Install one shape for every (kind, class) it is compiled for at one
    digit width."""
    installed = 0
    for k in KINDS:
        for c in CLASSES:
            try:
                sufkit_amd.tune_set([{"kind": k, "size_class": c, "block": block, "items": items, "chosen": 1,
                                      "digit_bits": digit_bits, "us_per_pass": us_per_pass}])
                installed += 1
            except sufkit_amd.SufkitError:
                pass
    return installed


def exact(hip, text, bits):
    sa, isa, lcp, _ = hip.hip_build(text, construction_bits=bits)
    osa, oisa, olcp = oracle.fast_build(text)
    assert np.array_equal(sa.astype(np.uint64), osa)
    assert np.array_equal(isa.astype(np.uint64), oisa)
    assert np.array_equal(lcp.astype(np.uint64), olcp)


@pytest.mark.gpu
@pytest.mark.parametrize("shape", SHAPES)
def test_every_shape_builds_exactly(hip, shape):
    sufkit_amd.tune_reset()
    try:
        assert install_everywhere(*shape) > 0
        text = synth.repeat_text(5, 300000, cover=0.5, n_runs=3, contigs=4)
        for bits in (32, 64):
            exact(hip, text, bits)
        texts = [synth.random_text(700 + r, 5000 + 3000 * r) for r in range(6)]
        for bits in (32, 64):
            parts, _ = hip.hip_build_batch(texts, construction_bits=bits)
            for t, (sa, isa, lcp) in zip(texts, parts):
                osa, oisa, olcp = oracle.fast_build(t)
                assert np.array_equal(sa.astype(np.uint64), osa) and np.array_equal(lcp.astype(np.uint64), olcp)
    finally:
        sufkit_amd.tune_reset()


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(512, 8), (1024, 16), (512, 24)])
def test_seven_bit_digits_build_exactly(hip, shape):
    """Times that favour 7-bit digits (1 us per 7-bit pass against 100 per
    8-bit one) make every keygen sort take them: builds stay exact."""
    sufkit_amd.tune_reset()
    try:
        assert install_everywhere(*shape, digit_bits=8, us_per_pass=100.0) > 0
        assert install_everywhere(*shape, digit_bits=7, us_per_pass=1.0) > 0
        text = synth.repeat_text(9, 300000, cover=0.5, n_runs=3, contigs=4)
        for bits in (32, 64):
            exact(hip, text, bits)
        exact(hip, synth.random_text(11, 200000), 64)
        texts = [synth.random_text(900 + r, 5000 + 3000 * r) for r in range(6)]
        for bits in (32, 64):
            parts, _ = hip.hip_build_batch(texts, construction_bits=bits)
            for t, (sa, isa, lcp) in zip(texts, parts):
                osa, oisa, olcp = oracle.fast_build(t)
                assert np.array_equal(sa.astype(np.uint64), osa) and np.array_equal(lcp.astype(np.uint64), olcp)
    finally:
        sufkit_amd.tune_reset()


@pytest.mark.gpu
def test_autotune_installs_the_fastest_measured_shape(hip):
    sufkit_amd.tune_reset()
    try:
        entries = sufkit_amd.autotune(1 << 20)
        assert entries
        by_cell = {}
        for e in entries:
            assert (e["block"], e["items"]) in SHAPES and e["us_per_pass"] > 0
            assert e["digit_bits"] == 8 or (e["digit_bits"] == 7 and e["kind"] < 4)
            by_cell.setdefault((e["kind"], e["size_class"], e["digit_bits"]), []).append(e)
        assert {k for k, _, _ in by_cell} == set(KINDS)
        assert {b for _, _, b in by_cell} == {7, 8}
        table = {(e["kind"], e["size_class"], e["digit_bits"]): e for e in sufkit_amd.tune_get()}
        for cell, measured in by_cell.items():
            chosen = [e for e in measured if e["chosen"]]
            assert len(chosen) == 1
            assert chosen[0]["us_per_pass"] == min(e["us_per_pass"] for e in measured)
            assert (table[cell]["block"], table[cell]["items"]) == (chosen[0]["block"], chosen[0]["items"])
            assert table[cell]["tuned"] == 1 and table[cell]["us_per_pass"] == chosen[0]["us_per_pass"]
        exact(hip, synth.repeat_text(6, 200000), 32)
        exact(hip, synth.repeat_text(7, 200000), 64)
    finally:
        sufkit_amd.tune_reset()

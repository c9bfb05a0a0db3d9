"""GPU parity: HIP SA/ISA/LCP vs the oracle, bit-exact, through the C ABI.

Mirrors the reference's backend differential tests (test_caps_backend.cpp:
43-94 exact-run texts vs an independent sort + scalar LCP oracle, :126-139
row-by-row SA comparison) and the fixture checks (test_sufkit.cpp:220-243).
"""
import numpy as np
import pytest

import oracle
from tests import synth

pytestmark = pytest.mark.gpu


def check_against_oracle(hip, text, construction_bits=32, method="auto", **kw):
    sa, isa, lcp, stats = hip.hip_build(text, construction_bits, **kw)
    osa, oisa, olcp = oracle.fast_build(text, method)
    n = text.size
    assert sa.size == n
    np.testing.assert_array_equal(sa.astype(np.uint64), osa)
    np.testing.assert_array_equal(isa.astype(np.uint64), oisa)
    np.testing.assert_array_equal(lcp.astype(np.uint64), olcp)
    return stats


@pytest.mark.parametrize("bits", [32, 64])
def test_legacy_fixture_vectors(hip, legacy_fixtures, bits):
    v = legacy_fixtures["legacy-sa-v1.1"]["vectors"]
    text = np.array(v["text"], dtype=np.uint8)
    np.testing.assert_array_equal(text, synth.fixture_text())
    sa, isa, lcp, _ = hip.hip_build(text, bits)
    assert sa.tolist() == v["sa"]
    assert isa.tolist() == v["isa"]
    assert lcp.tolist() == v["lcp"]


@pytest.mark.parametrize("length", [1, 2, 3, 5, 31, 32, 33, 63, 64, 65, 100, 257, 1000, 4097])
def test_random_small(hip, length):
    for seed in range(3):
        check_against_oracle(hip, synth.random_text(seed * 1000 + length, length), method="naive")


# 64-bit construction carries each suffix's second 16 symbols in its list
# values through the exact 32-symbol step and reads only windows near a
# separator, N or the sentinel: these texts put such windows everywhere.
@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("length", [1, 7, 40, 300, 2000, 20000])
def test_random_symbols(hip, length, bits):
    for seed in range(3):
        check_against_oracle(hip, synth.random_symbols(seed + length, length), bits)


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("run", [16, 31, 32, 33, 64, 100, 1000, 4096, 50000])
@pytest.mark.parametrize("kind", ["A", "N", "tandem2", "tandem7"])
def test_runs(hip, run, kind, bits):
    text = synth.run_text(run, [("rand", 50), (kind, run), ("rand", 20), ("sep", 1),
                                (kind, run // 2 + 1), ("sep", 1), ("rand", 10)])
    check_against_oracle(hip, text, bits)


def test_only_runs(hip):
    # Every suffix ties with many others for many doubling rounds.
    check_against_oracle(hip, synth.run_text(1, [("A", 70000)]))
    check_against_oracle(hip, synth.run_text(2, [("N", 70000)]))
    check_against_oracle(hip, synth.run_text(3, [("A", 3000), ("sep", 1), ("A", 3000), ("sep", 1), ("A", 2999)]))


@pytest.mark.parametrize("bits", [32, 64])
def test_long_run_one_segment_rounds(hip, bits):
    # The run's suffixes stay one group for many doubling rounds, so those
    # rounds sort and apply a one-segment list of more than 512 tiles: the
    # reduce / spine / apply schedule, with groups that run in from earlier
    # tiles and ranks that the round does not change.
    text = synth.run_text(7, [("rand", 300000), ("N", 2_600_000), ("rand", 200000)])
    stats = check_against_oracle(hip, text, bits)
    assert stats["doubling_rounds"] >= 15


@pytest.mark.parametrize("bits", [32, 64])
def test_repeats(hip, bits):
    text = synth.repeat_text(5, 200_000, cover=0.6, divergence=(0.0, 0.02), n_runs=10)
    check_against_oracle(hip, text, construction_bits=bits)


@pytest.mark.parametrize("name", ["random_1M", "repeats_600k", "runs_300k", "symbols_200k"])
@pytest.mark.parametrize("bits", [32, 64])
def test_reference_digests(hip, reference_digests, name, bits):
    text = dict(synth.golden_cases())[name]
    sa, isa, lcp, _ = hip.hip_build(text, bits)
    want = reference_digests[name]["ref32" if bits == 32 else "ref64"]
    assert f"0x{oracle.digest(sa.astype(np.uint64)):016x}" == want["sa"]
    assert f"0x{oracle.digest(isa.astype(np.uint64)):016x}" == want["isa"]
    assert f"0x{oracle.digest(lcp.astype(np.uint64)):016x}" == want["lcp"]


@pytest.mark.parametrize("widths", [(64, 64, 64), (32, 64, 32), (64, 32, 64)])
def test_output_widths(hip, widths):
    text = synth.repeat_text(9, 50_000)
    sa_bits, isa_bits, lcp_bits = widths
    for construction in (32, 64):
        check_against_oracle(hip, text, construction, sa_bits=sa_bits, isa_bits=isa_bits,
                             lcp_bits=lcp_bits)


def test_invalid_symbol_rejected(hip):
    text = synth.random_text(1, 100)
    text[5] = 9
    with pytest.raises(hip.SufkitError) as e:
        hip.hip_build(text)
    assert e.value.code == "invalid_input"


def test_missing_sentinel_rejected(hip):
    text = synth.random_text(1, 100)
    text[-1] = 2
    with pytest.raises(hip.SufkitError) as e:
        hip.hip_build(text)
    assert e.value.code == "invalid_input"
    text = synth.random_text(1, 100)
    text[10] = 0
    with pytest.raises(hip.SufkitError) as e:
        hip.hip_build(text)
    assert e.value.code == "invalid_input"


def test_sa_only(hip):
    text = synth.random_text(4, 5000)
    sa, isa, lcp, _ = hip.hip_build(text, retain_isa=False, retain_lcp=False)
    assert isa is None and lcp is None
    np.testing.assert_array_equal(sa.astype(np.uint64), oracle.sa_doubling(text))

"""Pins the CPU oracle before anything is compared against it (CPU only).

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong sign or index) fails it: the reference's own golden vectors,
brute force on small texts, closed forms, and the compiled reference.
"""
import numpy as np
import pytest

import oracle
from tests import synth


def vectors(fixtures, name):
    return fixtures[name]["vectors"]


def test_encode_records_matches_fixture_text(legacy_fixtures):
    # FromRecords of tests/data/reference.fa (genome_reference.cpp:131-160).
    text, amb = oracle.encode_records([b"ACGTNACGT", b"TTTACGTAAA", b"AAA"])
    assert text.tolist() == vectors(legacy_fixtures, "legacy-sa-v1.1")["text"]
    assert amb == legacy_fixtures["legacy-sa-v1.1"]["ambiguous_bases"] == 1
    # lowercase and IUPAC codes normalise like genome_reference.cpp:30-46
    lower, amb2 = oracle.encode_records([b"acgtRYn"])
    assert lower.tolist() == [2, 3, 4, 5, 6, 6, 6, 1, 0] and amb2 == 3


def test_fingerprint_matches_fixture(legacy_fixtures):
    text = np.array(vectors(legacy_fixtures, "legacy-sa-v1.1")["text"], np.uint8)
    fp = oracle.fingerprint(text[:-1])  # FNV over encoded bytes, sentinel excluded
    assert f"0x{fp:016x}" == legacy_fixtures["legacy-sa-v1.1"]["fingerprint"]
    assert oracle.fingerprint(np.zeros(0, np.uint8)) == 14695981039346656037


@pytest.mark.parametrize("method", ["naive", "doubling"])
def test_sa_isa_lcp_match_fixture(legacy_fixtures, method):
    v = vectors(legacy_fixtures, "legacy-sa-v1.1")
    text = np.array(v["text"], np.uint8)
    sa, isa, lcp = oracle.fast_build(text, method)
    assert sa.tolist() == v["sa"]
    assert isa.tolist() == v["isa"]
    assert lcp.tolist() == v["lcp"]


def test_sampled_kasai_matches_fixture(legacy_fixtures):
    # legacy-sa-v1.3: K = 2 sampled SA, dense ISA over pos/K, generalised
    # Kasai with carry common -= K (divsufsort_backend.cpp:63-150).
    v = vectors(legacy_fixtures, "legacy-sa-v1.3")
    k = v["sampling_rate"]
    text = np.array(v["text"], np.uint8)
    full = oracle.sa_naive(text)
    sampled = full[full % k == 0]
    assert sampled.tolist() == v["sa"]
    isa = oracle.isa(sampled, k)
    assert isa.tolist() == v["isa"]
    assert oracle.lcp_kasai(text, sampled, isa, k).tolist() == v["lcp"]


@pytest.mark.parametrize("seed", range(6))
def test_doubling_equals_naive(seed):
    for text in (synth.random_text(seed, 3000), synth.random_symbols(seed, 2000),
                 synth.run_text(seed, [("A", 500), ("rand", 50), ("tandem3", 400), ("sep", 1),
                                       ("N", 300), ("tandem7", 200)])):
        assert np.array_equal(oracle.sa_doubling(text), oracle.sa_naive(text))


def test_sa_is_sorted_permutation():
    text = synth.repeat_text(3, 5000, cover=0.5)
    sa = oracle.sa_doubling(text)
    assert np.array_equal(np.sort(sa), np.arange(text.size))
    suffixes = [bytes(text[i:]) for i in sa]
    assert all(a < b for a, b in zip(suffixes, suffixes[1:]))


@pytest.mark.parametrize("seed", range(4))
def test_kasai_equals_brute_force(seed):
    text = synth.run_text(seed, [("rand", 300), ("A", 200), ("rand", 40), ("tandem2", 300),
                                 ("sep", 1), ("N", 100), ("rand", 200)])
    sa = oracle.sa_naive(text)
    isa = oracle.isa(sa)
    assert np.array_equal(isa[sa], np.arange(text.size))
    assert np.array_equal(oracle.lcp_kasai(text, sa, isa), oracle.lcp_brute(text, sa))


def test_planted_lcp_mistakes_are_caught():
    # A wrong carry, an off-by-one row, or a missing LCP[0]=0 must all fail
    # the brute-force pin used above.
    text = synth.repeat_text(4, 3000, cover=0.6)
    sa = oracle.sa_naive(text)
    good = oracle.lcp_brute(text, sa)
    shifted = np.roll(good, 1)
    bumped = good.copy()
    bumped[len(bumped) // 2] += 1
    nonzero_head = good.copy()
    nonzero_head[0] = 1
    for bad in (shifted, bumped, nonzero_head):
        assert not np.array_equal(bad, good)


def test_crc32_check_value_and_fixture_crcs(legacy_fixtures):
    assert oracle.crc32(b"123456789") == 0xCBF43926
    assert oracle.crc32(b"6789", oracle.crc32(b"12345")) == 0xCBF43926
    for name in ("legacy-sa-v1.0", "legacy-sa-v1.1", "legacy-sa-v1.2", "legacy-sa-v1.3"):
        fx = legacy_fixtures[name]
        header = bytes.fromhex(fx["header_zeroed_crc_hex"])
        assert oracle.crc32(header) == fx["header_crc"]
        for section in fx["sections"]:
            assert oracle.crc32(bytes.fromhex(section["hex"])) == section["crc"]


@pytest.mark.skipif(not oracle.ref_available(), reason="compiled reference not built")
@pytest.mark.parametrize("method", ["ref32", "ref64"])
def test_compiled_reference_matches_fixture_and_oracle(legacy_fixtures, method):
    v = vectors(legacy_fixtures, "legacy-sa-v1.1")
    sa, isa, lcp = oracle.fast_build(np.array(v["text"], np.uint8), method)
    assert (sa.tolist(), isa.tolist(), lcp.tolist()) == (v["sa"], v["isa"], v["lcp"])
    text = synth.repeat_text(8, 20000, cover=0.5, n_runs=6)
    got = oracle.fast_build(text, method)
    want = oracle.fast_build(text, "doubling")
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("name", ["runs_300k", "symbols_200k"])
def test_oracle_matches_reference_digests(reference_digests, name):
    # Committed digests of the compiled reference (tests/golden/make_golden.py).
    text = dict(synth.golden_cases())[name]
    sa, isa, lcp = oracle.fast_build(text, "doubling")
    want = reference_digests[name]["ref32"]
    assert f"0x{oracle.digest(sa):016x}" == want["sa"]
    assert f"0x{oracle.digest(isa):016x}" == want["isa"]
    assert f"0x{oracle.digest(lcp):016x}" == want["lcp"]


@pytest.mark.skipif(not oracle.ref_caps_available(), reason="compiled reference CaPS-SA not built")
@pytest.mark.parametrize("bits", [32, 64])
def test_reference_caps_matches_the_oracle(bits):
    # The multi-core CPU baseline agrees with the restatement on ACGT,
    # 7-symbol and run-heavy texts.
    for text in (synth.random_text(41, 3000), synth.random_symbols(42, 2500),
                 synth.run_text(43, [("rand", 40), ("A", 700), ("tandem7", 400), ("sep", 1), ("N", 90)])):
        sa, isa_, lcp, _ = oracle.ref_caps_fast(text, bits, 4)
        osa, oisa, olcp = oracle.fast_build(text, "doubling")
        assert np.array_equal(sa, osa) and np.array_equal(isa_, oisa) and np.array_equal(lcp, olcp)


def test_sum_digest_sees_changed_swapped_and_shifted_values():
    # The digest the device reduction computes: a changed value, two swapped
    # rows and a shifted run each change it; the value width does not.
    values = np.arange(7, 40007, dtype=np.uint64) * 3
    base = oracle.sum_digest(values)
    changed = values.copy()
    changed[1234] += 1
    swapped = values.copy()
    swapped[[5, 30000]] = swapped[[30000, 5]]
    shifted = np.concatenate([values[1:], values[:1]])
    digests = {base, oracle.sum_digest(changed), oracle.sum_digest(swapped), oracle.sum_digest(shifted),
               oracle.sum_digest(values[:-1])}
    assert len(digests) == 5
    assert oracle.sum_digest(values.astype(np.uint32)) == base
    assert oracle.sum_digest(np.zeros(0, np.uint64)) == 0


def _lean_cases():
    return [
        ("random", synth.random_text(61, 20000)),
        ("symbols", synth.random_symbols(62, 15000)),
        ("runs", synth.run_text(63, [("rand", 300), ("A", 2500), ("tandem2", 900), ("N", 400), ("sep", 1),
                                     ("tandem7", 1200), ("rand", 500)])),
        ("repeats", synth.repeat_text(64, 30000)),
    ]


@pytest.mark.parametrize("name,text", _lean_cases())
def test_lean_digests_equal_the_full_oracle(name, text):
    # The lean digests (ISA scatter and Phi/PLCP from a 32-bit SA, the form
    # the 3.1 Gbp golden digests use) equal the digests of the oracle's
    # pinned ISA and Kasai LCP arrays, and LCP max.
    sa, isa_, lcp = oracle.fast_build(text, "doubling")
    got = oracle.lean_digests(text, sa.astype(np.uint32))
    assert got["sa"] == oracle.sum_digest(sa)
    assert got["isa"] == oracle.sum_digest(isa_)
    assert got["lcp"] == oracle.sum_digest(lcp)
    assert got["lcp_max"] == int(lcp.max())


@pytest.mark.skipif(not oracle.ref_available(), reason="compiled reference not built")
@pytest.mark.parametrize("name,text", _lean_cases()[2:])
def test_lean_digests_equal_the_compiled_reference(name, text):
    sa, isa_, lcp = oracle.fast_build(text, "ref64")
    got = oracle.lean_digests(text, sa.astype(np.uint32))
    assert (got["sa"], got["isa"], got["lcp"]) == tuple(oracle.sum_digest(a) for a in (sa, isa_, lcp))
    ref_sa, secs = oracle.ref_divsufsort64_u32(text)
    assert np.array_equal(ref_sa, sa.astype(np.uint32)) and secs >= 0


def test_lean_digests_catch_planted_mistakes():
    text = synth.repeat_text(65, 20000)
    sa, isa_, lcp = oracle.fast_build(text, "doubling")
    sa32 = sa.astype(np.uint32)
    want = oracle.lean_digests(text, sa32)
    # Two adjacent rows whose LCPs with their neighbours differ, so that the
    # swap changes the LCP as well as the order.
    r = int(np.flatnonzero((lcp[1:-1] != lcp[2:]) & (lcp[1:-1] > 0))[10]) + 1
    swapped = sa32.copy()
    swapped[[r, r + 1]] = swapped[[r + 1, r]]
    got = oracle.lean_digests(text, swapped)
    assert got["sa"] != want["sa"] and got["isa"] != want["isa"] and got["lcp"] != want["lcp"]
    repeated = sa32.copy()
    repeated[7] = repeated[8]
    with pytest.raises(RuntimeError):
        oracle.lean_digests(text, repeated)

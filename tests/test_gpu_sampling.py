"""Text-position sampling (K > 1) on the device, through the C ABI
(sfk_hip_build / sfk_hip_build_device with sampling_rate): the sampled SA,
the ISA over positions / K and the interval-minimum LCP, equal to the
reference's own BuildDivsufsort32/64 with the same K (compiled into
oracle/_ref), to the legacy K = 2 fixture and to the oracle's generalized
Kasai (divsufsort_backend.cpp:62-150)."""
import numpy as np
import pytest

import oracle
import sufkit_amd
from tests import synth

pytestmark = pytest.mark.gpu

TEXTS = {
    "random": lambda: synth.random_text(31, 30001),
    "symbols": lambda: synth.random_symbols(32, 7777),
    "runs": lambda: synth.run_text(33, [("rand", 50), ("A", 2500), ("sep", 1), ("tandem3", 700), ("N", 900),
                                        ("rand", 20)]),
    "repeats": lambda: synth.repeat_text(34, 40000, cover=0.6, n_runs=5, contigs=4),
}


def test_fixture_k2_through_the_c_abi(hip, legacy_fixtures):
    text = np.array(legacy_fixtures["legacy-sa-v1.1"]["vectors"]["text"], np.uint8)
    v = legacy_fixtures["legacy-sa-v1.3"]["vectors"]
    sa, isa, lcp, _ = hip.hip_build(text, 32, sampling_rate=2)
    assert sa.tolist() == v["sa"] and isa.tolist() == v["isa"] and lcp.tolist() == v["lcp"]


@pytest.mark.parametrize("name", sorted(TEXTS))
@pytest.mark.parametrize("k", [2, 3, 8, 64])
@pytest.mark.parametrize("bits", [32, 64])
def test_sampled_equals_reference(hip, name, k, bits):
    text = TEXTS[name]()
    rsa, risa, rlcp = oracle.ref_build_sampled(text, k, bits)
    sa, isa, lcp, _ = hip.hip_build(text, bits, sampling_rate=k)
    assert sa.size == (text.size + k - 1) // k
    assert np.array_equal(sa.astype(np.uint64), rsa)
    assert np.array_equal(isa.astype(np.uint64), risa)
    assert np.array_equal(lcp.astype(np.uint64), rlcp)
    # the oracle's generalized Kasai on the filtered complete SA agrees
    full = oracle.sa_doubling(text)
    sampled = full[full % k == 0]
    assert np.array_equal(rsa, sampled)
    assert np.array_equal(rlcp, oracle.lcp_kasai(text, sampled, oracle.isa(sampled, k), k))


@pytest.mark.parametrize("k", [2, 5])
def test_sampled_device_entry_mixed_widths(hip, k):
    text = TEXTS["repeats"]()
    n = text.size
    rows = (n + k - 1) // k
    rsa, risa, rlcp = oracle.ref_build_sampled(text, k)
    d_text = sufkit_amd.DeviceArray(n + 64)
    d_text.upload(text)
    d_sa, d_isa, d_lcp = (sufkit_amd.DeviceArray(rows * 8) for _ in range(3))
    with sufkit_amd.DeviceBuilder(0) as b:
        b.build(d_text.ptr, n, 64, d_sa.ptr, 32, d_isa.ptr, 64, d_lcp.ptr, 32, sampling_rate=k)
        # SA only (no ISA, no LCP)
        d_sa2 = sufkit_amd.DeviceArray(rows * 8)
        b.build(d_text.ptr, n, 32, d_sa2.ptr, 64, None, 32, None, 32, flags=0, sampling_rate=k)
    assert np.array_equal(d_sa.download(np.uint32, rows).astype(np.uint64), rsa)
    assert np.array_equal(d_isa.download(np.uint64, rows), risa)
    assert np.array_equal(d_lcp.download(np.uint32, rows).astype(np.uint64), rlcp)
    assert np.array_equal(d_sa2.download(np.uint64, rows), rsa)


def test_sampling_rate_larger_than_text(hip):
    text = synth.random_text(35, 100)
    sa, isa, lcp, _ = hip.hip_build(text, 32, sampling_rate=1000)
    assert sa.tolist() == [0] and isa.tolist() == [0] and lcp.tolist() == [0]


def test_zero_sampling_rate_rejected(hip):
    with pytest.raises(sufkit_amd.SufkitError) as e:
        hip.hip_build(synth.random_text(36, 100), 32, sampling_rate=0)
    assert e.value.status == 1

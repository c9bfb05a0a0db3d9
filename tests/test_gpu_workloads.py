"""Every benchmark workload at its full size on the device, against the
compiled reference's own results.

tests/golden/workload_digests.json holds, per workload text (the
deterministic generator of sufkit_amd/host/synthetic.cpp), the text's
FNV-1a-64 fingerprint and the sum digests of the SA, ISA and LCP that the
reference's divsufsort construction computed in the build container
(tests/golden/make_workload_digests.py).  Here the device builds the same
text, the device digests its arrays (sfk_hip_digest) and the device checker
(sfk_hip_check_index) proves order, inverse and LCP row by row, so a result
is bit-exact with the reference at the full size, not only on samples.

  4 Mbp i.i.d. genome            32- and 64-bit construction
  64 Mbp multi-contig genome     32- and 64-bit construction
  512 bacterial references       one batched build, every reference's digests
  3.1 Gbp human-scale genome     64-bit construction, 32-bit storage
"""
import json
import sys

import numpy as np
import pytest

import oracle
import sufkit_amd
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

SEED = 20260822
FLAGS = sufkit_amd.SFK_HIP_RETAIN_ISA | sufkit_amd.SFK_HIP_RETAIN_LCP


@pytest.fixture(scope="module")
def digests():
    return json.loads((GOLDEN / "workload_digests.json").read_text())


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def _hex(v):
    return f"0x{v:016x}"


def build_and_check(text, bits, want, out_bits=32):
    """Device build of a host text, then device digests and the full check."""
    n = int(text.size)
    assert want["n"] == n
    assert want["fingerprint"] == _hex(oracle.fingerprint(text[:-1])), "generator produced another text"
    ob = out_bits // 8
    d_text = sufkit_amd.DeviceArray(n + 64)
    d_text.upload(text)
    outs = [sufkit_amd.DeviceArray(n * ob) for _ in range(3)]
    with sufkit_amd.DeviceBuilder(0) as builder:
        stats = builder.build(d_text.ptr, n, bits, outs[0].ptr, out_bits, outs[1].ptr, out_bits,
                              outs[2].ptr, out_bits, FLAGS)
    got = {k: _hex(sufkit_amd.device_digest(a.ptr, n, out_bits)) for k, a in zip(("sa", "isa", "lcp"), outs)}
    report = sufkit_amd.check_index(d_text.ptr, n, outs[0].ptr, out_bits, outs[1].ptr, out_bits,
                                    outs[2].ptr, out_bits, raise_on_failure=False)
    _say(f"n={n} bits={bits} device {stats['device_seconds']:.3f} s, check {report['seconds']:.2f} s")
    for a in (d_text, *outs):
        a.free()
    return got, report


@pytest.mark.parametrize("bits", [32, 64])
def test_4mbp_iid_genome(hip, digests, bits):
    text, _ = sufkit_amd.synthetic_text(1, SEED)
    got, report = build_and_check(text, bits, digests["c1"])
    assert report["status"] == 0 and report["lcp_checked"] == 1, report
    assert got == {k: digests["c1"][k] for k in ("sa", "isa", "lcp")}


@pytest.mark.parametrize("bits", [32, 64])
def test_64mbp_multicontig_genome(hip, digests, bits):
    text, _ = sufkit_amd.synthetic_text(2, SEED)
    got, report = build_and_check(text, bits, digests["c2"])
    assert report["status"] == 0, report
    assert got == {k: digests["c2"][k] for k in ("sa", "isa", "lcp")}


@pytest.mark.timeout(900)
def test_3100mbp_human_scale_genome(hip, digests):
    want = digests.get("c3")
    if want is None:
        pytest.skip("3.1 Gbp reference digests not generated")
    text, _ = sufkit_amd.synthetic_text(3, SEED, threads=16)
    got, report = build_and_check(text, 64, want)
    assert report["status"] == 0 and report["rows"] == text.size, report
    assert got == {k: want[k] for k in ("sa", "isa", "lcp")}


@pytest.mark.timeout(900)
def test_512_bacterial_references_batched(hip, digests):
    want = digests["c4"]["per_reference"]
    texts = [sufkit_amd.synthetic_text(4, SEED, r)[0] for r in range(len(want))]
    sizes = np.array([t.size for t in texts], np.uint64)
    offsets = np.zeros(len(texts) + 1, np.uint64)
    offsets[1:] = np.cumsum(sizes)
    n = int(offsets[-1])
    d_text = sufkit_amd.DeviceArray(n + 64)
    d_text.upload(np.concatenate(texts))
    outs = [sufkit_amd.DeviceArray(n * 4) for _ in range(3)]
    with sufkit_amd.DeviceBuilder(0) as builder:
        stats = builder.build_batch(d_text.ptr, n, offsets, 32, outs[0].ptr, 32, outs[1].ptr, 32,
                                    outs[2].ptr, 32, FLAGS)
    _say(f"batch of {len(texts)}: n={n} device {stats['device_seconds']:.3f} s")
    bad = []
    for r, (t, w) in enumerate(zip(texts, want)):
        a, m = int(offsets[r]), int(sizes[r])
        assert w["n"] == m and w["fingerprint"] == _hex(oracle.fingerprint(t[:-1]))
        got = {k: _hex(sufkit_amd.device_digest(arr.ptr + 4 * a, m, 32))
               for k, arr in zip(("sa", "isa", "lcp"), outs)}
        if got != {k: w[k] for k in ("sa", "isa", "lcp")}:
            bad.append(r)
        if r % 64 == 0:  # the full check on every 64th reference as well
            rep = sufkit_amd.check_index(d_text.ptr + a, m, outs[0].ptr + 4 * a, 32, outs[1].ptr + 4 * a, 32,
                                         outs[2].ptr + 4 * a, 32, raise_on_failure=False)
            assert rep["status"] == 0, (r, rep)
    for a in (d_text, *outs):
        a.free()
    assert bad == [], f"references whose digests differ from the reference's: {bad[:10]}"

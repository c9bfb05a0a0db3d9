#!/usr/bin/env python3
"""Benchmark: SA + ISA + raw LCP construction (Fast profile) on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 3]

A step is one complete device build (pack, prefix doubling, ISA validation,
Φ/PLCP/LCP) of the workload's logical text, already resident in HBM, into
device SA/ISA/LCP buffers.  The default workload is the 3.1 Gbp human-scale
genome (24 contigs of 50-250 Mbp, GC 41%, 45% repeats, 3% satellites, N gaps
up to 10 Mbp) with 64-bit construction and validated 32-bit storage.

Prints one JSON line.  `value` is bases / 1e6 / seconds-per-step.  The
`roofline` object is for the dominant kernel (the LSD radix scatter pass),
timed with HIP events on the build stream during the timed steps.  After
the timed steps the device digests its SA/ISA/LCP and compares them with the
reference's own digests of the same text (tests/golden/workload_digests.json),
and the device checker proves order, inverse and LCP of every row.
`cpu_baseline` is the reference's own divsufsort backend (SA + ISA + Kasai
LCP, 1 thread; divsufsort64 for 64-bit construction) timed on this host on a
bounded prefix of the same text, with CaPS-SA on all host threads beside it.
`host_build` times the whole host SuffixArray::Build (host text to validated
host storage plus the Fast prefix directory), the reference's build_seconds
scope.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBPS = 8000.0

WORKLOADS = {
    1: "4 Mbp single contig, i.i.d. uniform A/C/G/T",
    2: ("64 Mbp multi-contig reference: 200 log-normal contigs (median 200 kbp), GC 45%, "
        "1% N runs of 10-5000, 30% repeats (300 bp-6 kbp families, 1-15% substitutions), "
        "1 Mbp poly-A/period-2/period-7 block"),
    3: ("3.1 Gbp human-scale genome: 24 contigs of 50-250 Mbp, GC 41%, 45% repeats, "
        "3% satellites, N gaps up to 10 Mbp"),
    4: ("batched build: 512 bacterial-like references, log-normal 0.5-12 Mbp (median 4 Mbp), "
        "GC 30-70%, 5% shared mobile elements; one batched device build vs one build per reference"),
    5: ("served build loop: Poisson arrivals of references log-uniform 1-200 Mbp (64 Mbp-workload "
        "generator), admission by HBM footprint, indexes resident, 1e6 exact count queries per "
        "index (length 20-32, 50% sampled from the text, 50% random)"),
}


def parse_args():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", type=int, default=3, choices=sorted(WORKLOADS))
    ap.add_argument("--scale", type=float, default=1.0, help="workload size multiplier")
    ap.add_argument("--construction-bits", type=int, default=None, choices=(32, 64))
    ap.add_argument("--seed", type=int, default=20260822)
    ap.add_argument("--refs", type=int, default=512, help="references in the batched workload")
    ap.add_argument("--requests", type=int, default=16, help="requests in the served workload")
    ap.add_argument("--arrival-ms", type=float, default=20.0, help="mean inter-arrival time (served)")
    ap.add_argument("--slots", type=int, default=4, help="concurrent builds (served)")
    ap.add_argument("--queries", type=int, default=1_000_000, help="count queries per index (served)")
    ap.add_argument("--sa-count", action="store_true",
                    help="served: count by SA binary search instead of FM backward search")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-host-build", action="store_true",
                    help="skip the timed host SuffixArray::Build leg")
    ap.add_argument("--no-extra-leg", action="store_true",
                    help="skip the 64 Mbp 32-bit leg that the 3.1 Gbp run adds")
    ap.add_argument("--cpu-sample-symbols", type=int, default=0,
                    help="symbols of the text timed on the CPU (0 = whole text)")
    ap.add_argument("--no-profile", action="store_true",
                    help="time without per-kernel HIP events (no roofline)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first run one device-to-device copy of this many MiB (counter calibration)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the built-in radix shapes instead of tuning them on the device first")
    ap.add_argument("--save", action="store_true",
                    help="also write the built index to a .sufidx file straight from the device arrays "
                         "(SaveDeviceIndex) and time the reference's own Save of the same index")
    ap.add_argument("--save-dir", default=None, help="directory for the --save files (default: a temporary one)")
    ap.add_argument("--low-memory", action="store_true",
                    help="also time the Low-memory profile's byte-coded LCP (sfk_hip_byte_coded_lcp) of the built "
                         "index and check it decodes to the raw LCP")
    ap.add_argument("--save-tuning", metavar="JSON", help="write the tuned radix shape table here")
    ap.add_argument("--load-tuning", metavar="JSON",
                    help="install a saved radix shape table instead of measuring one")
    return ap.parse_args()


def radix_tuning(args, max_elements: int):
    """Radix shapes for this run, outside the timed region: measured on the
    device by sfk_hip_autotune (default), loaded from a saved table, or the
    built-in ones.  Returns the summary reported in the JSON line."""
    import sufkit_amd

    sufkit_amd.tune_reset()
    if args.load_tuning:
        raw = Path(args.load_tuning).read_text()
        if args.load_tuning.endswith(".md"):
            # a profile summary holding the table in a ```json block
            raw = raw.split("```json", 1)[1].split("```", 1)[0]
        table = json.loads(raw)
        sufkit_amd.tune_set(table)
        source, seconds = "loaded", 0.0
    elif args.no_autotune:
        return {"source": "built-in"}
    else:
        t0 = time.perf_counter()
        measured = sufkit_amd.autotune(max_elements)
        seconds = time.perf_counter() - t0
        source = "sfk_hip_autotune"
        if args.save_tuning:
            Path(args.save_tuning).write_text(json.dumps([e for e in measured if e["chosen"]]))
    # per kind and class: the shape of each digit width with its time per pass
    chosen = {}
    for e in sufkit_amd.tune_get():
        if e["tuned"]:
            cell = chosen.setdefault(sufkit_amd.TUNE_KINDS[e["kind"]], {}).setdefault(f"class{e['size_class']}", [])
            cell.append(f"{e['digit_bits']}b {e['block']}x{e['items']} {e['us_per_pass']:.1f}us")
    return {"source": source, "seconds": round(seconds, 3), "max_elements": max_elements, "chosen": chosen}


# HBM traffic per RadixOnesweep launch, measured by two rocprofv3 --pmc runs
# (FETCH_SIZE, WRITE_SIZE; calibrated against a device copy of known size)
# and committed with their commands, per (workload, construction bits).
TRAFFIC_PROFILES = {
    (3, 64): ROOT / "profiles" / "sa_build_3100mbp_pmc_traffic.md",
    (2, 32): ROOT / "profiles" / "sa_build_64mbp_pmc_traffic.md",
}

GOLDEN_DIGESTS = ROOT / "tests" / "golden" / "workload_digests.json"


def measured_traffic(workload: int, bits: int) -> dict:
    path = TRAFFIC_PROFILES.get((workload, bits))
    if path is None:
        return {}
    try:
        for line in path.read_text().splitlines():
            if line.startswith("onesweep_traffic_bytes_per_launch:"):
                return {"traffic": int(line.split(":", 1)[1]),
                        "traffic_source": str(path.relative_to(ROOT))}
    except OSError:
        pass
    return {}


def golden_digests(workload: int, seed: int, scale: float) -> dict | None:
    """The reference's SA/ISA/LCP digests of this workload's text, when the
    golden file holds them (workloads 1-3 at the default seed and scale)."""
    key = {1: "c1", 2: "c2", 3: "c3"}.get(workload)
    if key is None or seed != 20260822 or scale != 1.0:
        return None
    try:
        return json.loads(GOLDEN_DIGESTS.read_text()).get(key)
    except (OSError, ValueError):
        return None


def device_parity(d_text, n, d_sa, d_isa, d_lcp, out_bits, want, fingerprint) -> dict:
    """Device digests of the built arrays against the reference's digests,
    and the device's full order/inverse/LCP check of every row."""
    import sufkit_amd

    got = {k: f"0x{sufkit_amd.device_digest(ptr, n, out_bits):016x}"
           for k, ptr in (("sa", d_sa), ("isa", d_isa), ("lcp", d_lcp))}
    out = {"digests": got}
    if want is not None:
        same = want.get("fingerprint") == fingerprint and all(got[k] == want[k] for k in got)
        out["parity_vs_reference_digests"] = "bit-exact" if same else "MISMATCH"
        out["reference_digest_method"] = want.get("method")
    rep = sufkit_amd.check_index(d_text, n, d_sa, out_bits, d_isa, out_bits, d_lcp, out_bits,
                                 raise_on_failure=False)
    out["full_check"] = {"rows": rep["rows"], "failed": rep["status"] != 0,
                         "order_failures": rep["order_failures"], "inverse_failures": rep["inverse_failures"],
                         "range_failures": rep["range_failures"],
                         "lcp_failures": rep["lcp_failures"] + rep["lcp_prefix_failures"],
                         "seconds": round(rep["seconds"], 3)}
    return out


def cpu_info() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(text: np.ndarray, sample_symbols: int, width_bits: int = 32):
    """Reference divsufsort32 Fast build (SA+ISA+LCP, 1 thread) on a bounded
    sample of the same text, or the oracle restatement when the compiled
    reference is absent."""
    import oracle

    n = text.size
    if sample_symbols and sample_symbols < n:
        sample = np.concatenate([text[: sample_symbols - 1], np.zeros(1, np.uint8)])
        # A prefix may cut inside a contig: keep it a valid logical text.
        sample[sample == 0] = 1
        sample[-1] = 0
    else:
        sample = text
    bases = int(np.count_nonzero(sample > 1))
    if oracle.ref_available():
        kind = "reference"
        method = "ref32" if width_bits == 32 and sample.size <= 0x7FFFFFFF else "ref64"
    else:
        kind = "port"
        method = "doubling"
    t0 = time.perf_counter()
    sa, isa, lcp = oracle.fast_build(sample, method)
    dt = time.perf_counter() - t0
    return {
        "value": round(bases / 1e6 / dt, 3), "unit": "Mbp/s", "cores": 1, "kind": kind,
        "sample": (f"{'whole' if sample is text else 'prefix of the'} workload text, "
                   f"{sample.size} symbols ({bases} bases); "
                   f"{ {'ref32': 'divsufsort32', 'ref64': 'divsufsort64'}.get(method, method)} "
                   "SA + ISA + Kasai LCP "
                   f"(detail::BuildDivsufsort, 1 thread) on {cpu_info()}"),
        "seconds": round(dt, 3),
    }, (sa, isa, lcp, sample is text, sample)


def host_threads() -> int:
    """Host cores this process may use, capped at the 16 a one-GPU box
    allots."""
    try:
        cores = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        cores = os.cpu_count() or 1
    return max(1, min(cores, 16))


def cpu_baseline_caps(sample: np.ndarray, width_bits: int, whole: bool):
    """The reference's CaPS-SA backend (SA + merge LCP) plus a threaded ISA
    scatter, on all host threads, on the same sample as the divsufsort
    baseline."""
    import oracle

    threads = host_threads()
    bases = int(np.count_nonzero(sample > 1))
    t0 = time.perf_counter()
    sa, isa, lcp, (t_sa_lcp, t_isa) = oracle.ref_caps_fast(sample, width_bits, threads)
    dt = time.perf_counter() - t0
    return {
        "value": round(bases / 1e6 / dt, 3), "unit": "Mbp/s", "cores": threads, "kind": "reference",
        "sample": (f"{'whole' if whole else 'prefix of the'} workload text, {sample.size} symbols ({bases} bases); "
                   f"CaPS-SA {width_bits}-bit SA + merge LCP (detail::BuildCaps{width_bits}) + ISA scatter, "
                   f"{threads} threads on {cpu_info()}"),
        "seconds": round(dt, 3), "sa_lcp_seconds": round(t_sa_lcp, 3), "isa_seconds": round(t_isa, 3),
    }, (sa, isa, lcp)


def run_batched(args) -> int:
    """Workload 4: the batched build against plain per-reference builds."""
    import sufkit_amd

    refs = args.refs
    texts, bases = [], 0
    for r in range(refs):
        t, lens = sufkit_amd.synthetic_text(4, args.seed, r, args.scale)
        texts.append(t)
        bases += int(lens.sum())
    offsets = np.zeros(refs + 1, dtype=np.uint64)
    offsets[1:] = np.cumsum([t.size for t in texts])
    n = int(offsets[-1])
    d_text = sufkit_amd.DeviceArray(n + 64)
    d_text.upload(np.concatenate(texts))
    del texts
    outs = {k: [sufkit_amd.DeviceArray(n * 4) for _ in range(3)] for k in ("batch", "plain")}
    tuning = radix_tuning(args, n)
    builder = sufkit_amd.DeviceBuilder(0)
    flags = sufkit_amd.SFK_HIP_RETAIN_ISA | sufkit_amd.SFK_HIP_RETAIN_LCP

    def batched():
        sa, isa, lcp = outs["batch"]
        return builder.build_batch(d_text.ptr, n, offsets, 32, sa.ptr, 32, isa.ptr, 32, lcp.ptr, 32, flags)

    def plain():
        sa, isa, lcp = outs["plain"]
        for a, b in zip(offsets[:-1].tolist(), offsets[1:].tolist()):
            builder.build(d_text.ptr + a, b - a, 32, sa.ptr + 4 * a, 32, isa.ptr + 4 * a, 32,
                          lcp.ptr + 4 * a, 32, flags)

    for _ in range(args.warmup):
        batched()
        plain()
    t0 = time.perf_counter()
    stats = [batched() for _ in range(args.steps)]
    t_batch = (time.perf_counter() - t0) / args.steps
    t0 = time.perf_counter()
    for _ in range(max(1, args.steps // 2)):
        plain()
    t_plain = (time.perf_counter() - t0) / max(1, args.steps // 2)
    identical = True
    chunk = 1 << 26
    for i in range(3):
        for start in range(0, n, chunk):
            cnt = min(chunk, n - start)
            if not np.array_equal(outs["batch"][i].download(np.uint32, cnt, start),
                                  outs["plain"][i].download(np.uint32, cnt, start)):
                identical = False
    last = stats[-1]
    result = {
        "metric": "SA+ISA+LCP build Mbp/s + HBM roofline frac, 1 MI355X vs divsufsort64 on host",
        "value": round(bases / 1e6 / t_batch, 2), "unit": "Mbp/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t_batch * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (deterministic splitmix64 generator, seed %d)" % args.seed,
        "config": {"workload": WORKLOADS[4], "references": refs, "text_symbols": n, "bases": bases},
        "plain_per_reference_mbps": round(bases / 1e6 / t_plain, 2),
        "plain_per_reference_ms": round(t_plain * 1e3, 3),
        "batched_over_plain": round(t_plain / t_batch, 3),
        "outputs_identical": identical,
        "phases": {k: last[k] for k in ("refine_steps", "doubling_rounds", "active_total", "sort_passes")},
        "radix_tuning": tuning,
    }
    builder.close()
    print(json.dumps(result), flush=True)
    return 0 if identical else 1


def run_served(args) -> int:
    """Workload 5: the serving loop (sufkit_amd/host/serve.cpp)."""
    import sufkit_amd

    rng = np.random.default_rng(args.seed)
    texts, bases = [], 0
    for r in range(args.requests):
        t, lens = sufkit_amd.synthetic_text(5, args.seed, r, args.scale)
        texts.append(t)
        bases += int(lens.sum())
    arrivals = np.cumsum(rng.exponential(args.arrival_ms / 1e3, size=args.requests)).tolist()
    tuning = radix_tuning(args, max(t.size for t in texts))
    # warm-up: one small request through the same loop (first-use costs)
    sufkit_amd.serve([sufkit_amd.synthetic_text(1, 1, 0, 0.01)[0]], [0.0],
                     slots=1, queries_per_index=1000)
    t0 = time.perf_counter()
    res, summary, samples = sufkit_amd.serve(texts, arrivals, slots=args.slots,
                                             queries_per_index=args.queries, sample_queries=100,
                                             fm_queries=not args.sa_count)
    wall = time.perf_counter() - t0
    # Check sampled counts of the smallest requests against a scan of the text.
    order = np.argsort([t.size for t in texts])[:3]
    bad = 0
    for i in order.tolist():
        s = "".join("$|ACGTN"[c] for c in texts[i].tolist())
        pats, lens_, cnts = samples[0][i], samples[1][i], samples[2][i]
        for k, (code, length, got) in enumerate(zip(pats, lens_, cnts)):
            p = sufkit_amd.decode_pattern(code, length)
            want, start = 0, s.find(p)
            while start != -1:
                want += 1
                start = s.find(p, start + 1)
            if want != got:
                bad += 1
                print(f"served count mismatch: request {i} (n {texts[i].size}, slot {res[i]['slot']}) "
                      f"query {k} {p} counted {got}, text scan {want}", file=sys.stderr, flush=True)
    lat = [r["built"] - r["arrival"] for r in res]
    result = {
        "metric": "SA+ISA+LCP build Mbp/s + HBM roofline frac, 1 MI355X vs divsufsort64 on host",
        "value": round(summary["build_mbps"], 2), "unit": "Mbp/s", "n_gpus": 1,
        "steps": 1, "warmup": 1, "ms_per_step": round(summary["makespan_seconds"] * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (deterministic splitmix64 generator, seed %d)" % args.seed,
        "config": {"workload": WORKLOADS[5], "requests": args.requests, "bases": bases,
                   "mean_interarrival_ms": args.arrival_ms, "slots": args.slots,
                   "queries_per_index": args.queries,
                   "count_queries": "SA binary search" if args.sa_count else "FM backward search"},
        "latency_p50_s": round(summary["latency_p50"], 4), "latency_p99_s": round(summary["latency_p99"], 4),
        "end_to_end_p50_s": round(summary["end_to_end_p50"], 4),
        "end_to_end_p99_s": round(summary["end_to_end_p99"], 4),
        "queries_per_second": round(summary["queries_per_second"], 1),
        "max_concurrent_builds": summary["max_concurrent"],
        "peak_reserved_GB": round(summary["peak_reserved_bytes"] / 1e9, 2),
        "requests_failed": summary["requests_failed"], "sampled_count_mismatches": bad,
        "first_error": (summary["first_error"].decode(errors="replace")
                        if isinstance(summary["first_error"], bytes) else summary["first_error"]) or None,
        "wall_seconds": round(wall, 3), "latency_max_s": round(max(lat), 4),
        "radix_tuning": tuning,
    }
    print(json.dumps(result), flush=True)
    return 0 if bad == 0 and summary["requests_failed"] == 0 else 1


def section_table(path: str) -> tuple[bytes, dict]:
    """Header bytes and {section type: (offset, size, crc)} of a .sufidx file."""
    import struct
    with open(path, "rb") as f:
        head = f.read(80)
        count, header_size = struct.unpack_from("<II", head, 24)
        head += f.read(header_size - 80)
    table = {}
    for i in range(count):
        stype, _, offset, size, crc, _ = struct.unpack_from("<IIQQII", head, 80 + 32 * i)
        table[stype] = (offset, size, crc)
    return head, table


def save_leg(args, text, contigs, d_text, d_sa, d_isa, d_lcp, bits, out_bits) -> dict:
    """Write the built index from the device arrays (SaveDeviceIndex: planes
    checksummed on the device, streamed through pinned buffers, the file
    reread and checked before it is published), then the reference's own
    SuffixArray::Build + Save of the same reference on the host: the saved
    sections must be identical, and the reference's Save time is the CPU
    baseline of this leg."""
    import shutil
    import tempfile
    from sufkit_amd import api
    import oracle
    ref = api.GenomeReference.from_encoded(text, contigs)
    if not np.array_equal(ref.text(), text):
        raise RuntimeError("the reference's normalized text differs from the benchmark text")
    opts = api.BuildOptions(backend=api.BACKEND_HIP, coordinate_width=bits)
    tmp = tempfile.mkdtemp(prefix="sufkit_save_", dir=args.save_dir)
    out: dict = {}
    try:
        # The first save on a device also sets up its plane writer (pinned
        # buffers); it is reported, and the second save is the timed one.
        ours = os.path.join(tmp, "device.sufidx")
        first = api.save_device_arrays(ref, ours, d_text, d_sa, out_bits, d_isa, out_bits, d_lcp, out_bits,
                                       bits, opts)
        os.remove(ours)
        st = api.save_device_arrays(ref, ours, d_text, d_sa, out_bits, d_isa, out_bits, d_lcp, out_bits, bits,
                                    opts)
        out = {
            "what": ("SaveDeviceIndex: .sufidx 1.4 from the device arrays (sections 1/2/3/5/6, CRC32 on the "
                     "device, pinned double buffering), reread with every CRC checked, then published"),
            "bytes": st["bytes"], "seconds": round(st["seconds"], 4),
            "write_seconds": round(st["write_seconds"], 4), "validate_seconds": round(st["validate_seconds"], 4),
            "GBps": round(st["bytes"] / st["seconds"] / 1e9, 3),
            "write_GBps": round(st["bytes"] / st["write_seconds"] / 1e9, 3),
            "first_save_seconds": round(first["seconds"], 4),
        }
        if oracle.ref_writer_available() and not args.no_cpu_baseline:
            letters = np.frombuffer(b"??ACGTN", np.uint8)
            records, at = [], 0
            for i, length in enumerate(contigs.tolist()):
                records.append((f"contig{i}", letters[text[at:at + length]].tobytes()))
                at += length + 1
            theirs = os.path.join(tmp, "reference.sufidx")
            t = oracle.ref_write_sufidx_timed(records, theirs, coordinate_width=bits)
            out["cpu_baseline"] = {
                "value": round(st["bytes"] / t["save_seconds"] / 1e9, 3), "unit": "GB/s", "cores": 1,
                "kind": "reference", "seconds": round(t["save_seconds"], 3),
                "sample": ("the same index: the reference's own SuffixArray::Save (serialization.cpp "
                           "WriteContainer, zlib CRC32 while writing, reread for validation) after its "
                           "divsufsort Build of the same records (not timed)"),
            }
            h_ours, s_ours = section_table(ours)
            h_theirs, s_theirs = section_table(theirs)
            same = (s_ours.keys() == s_theirs.keys() and all(s_ours[k] == s_theirs[k] for k in s_ours)
                    and os.path.getsize(ours) == os.path.getsize(theirs)
                    and {i for i in range(len(h_ours)) if h_ours[i] != h_theirs[i]} <= {14, 72, 73, 74, 75})
            out["parity_vs_reference_save"] = "sections identical" if same else "MISMATCH"
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return out


def low_memory_leg(hip, d_sa, d_isa, d_lcp, n, out_bits, sa, lcp) -> dict:
    """Byte-code the built LCP as the Low-memory profile stores it (primary
    byte per row, overflow anchors in text order), timed end to end into
    host memory, then decode every row back through the anchors and compare
    with the raw LCP when the full arrays are on the host."""
    hip.sfk_hip_byte_coded_lcp.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                           ctypes.c_uint32, ctypes.c_uint64, ctypes.c_void_p,
                                           ctypes.POINTER(ctypes.c_uint64), ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_uint64, ctypes.c_char_p, ctypes.c_size_t]
    hip.sfk_hip_byte_coded_lcp.restype = ctypes.c_int
    err = ctypes.create_string_buffer(512)
    primary = np.empty(n, np.uint8)
    count = ctypes.c_uint64()
    def run():
        if hip.sfk_hip_byte_coded_lcp(0, d_lcp, out_bits, d_isa, out_bits, n, primary.ctypes.data,
                                      ctypes.byref(count), None, None, 0, err, len(err)) != 0:
            raise RuntimeError(err.value.decode())
        pos = np.empty(count.value, np.uint64)
        val = np.empty(count.value, np.uint64)
        if count.value and hip.sfk_hip_byte_coded_lcp(0, d_lcp, out_bits, d_isa, out_bits, n, None,
                                                      ctypes.byref(count), pos.ctypes.data, val.ctypes.data,
                                                      pos.size, err, len(err)) != 0:
            raise RuntimeError(err.value.decode())
        return pos, val
    run()  # warm-up
    t0 = time.perf_counter()
    pos, val = run()
    seconds = time.perf_counter() - t0
    out = {"seconds": round(seconds, 4), "anchors": int(count.value),
           "rows_255_or_more": int((primary == 255).sum()),
           "bytes": int(n + 2 * pos.size * (4 if out_bits == 32 else 8)),
           "raw_lcp_bytes": int(n * out_bits // 8)}
    if sa is not None and lcp is not None:
        # decode: rows below 255 as they are; the others from the last anchor
        # at or before their suffix position
        rows = np.nonzero(primary == 255)[0]
        p = sa[rows].astype(np.uint64)
        k = np.searchsorted(pos, p, side="right") - 1
        decoded = primary.astype(np.uint64)
        ok = bool((k >= 0).all())
        if ok and rows.size:
            decoded[rows] = val[k] - (p - pos[k])
        out["decodes_to_raw_lcp"] = ok and bool(np.array_equal(decoded, lcp.astype(np.uint64)))
    return out


def host_build_leg(text: np.ndarray, contigs: np.ndarray, bits: int) -> dict:
    """The whole host SuffixArray::Build with the HIP backend, timed as the
    reference times build_seconds (sa_build_benchmark.cpp:548-553): host
    text to validated host SA/ISA/LCP storage plus the Fast prefix
    directory; the reference model is made before the clock starts."""
    from sufkit_amd import api

    bases = int(contigs.sum())
    ref = api.GenomeReference.from_encoded(text, contigs)
    opts = api.BuildOptions(backend=api.BACKEND_HIP, coordinate_width=bits)
    t0 = time.perf_counter()
    index = api.SuffixArray.build(ref, opts)
    wall = time.perf_counter() - t0
    st = index.stats
    info = index.info()
    out = {
        "t_build_seconds": round(wall, 3), "mbps": round(bases / 1e6 / wall, 2),
        "backend_call_seconds": round(st["call_seconds"], 3),
        "device_seconds": round(st["device_seconds"], 3),
        "pcie_seconds": round(st["transfer_seconds"], 3),
        "prefix_directory_seconds": round(st["prefix_directory_seconds"], 4),
        "host_storage_seconds": round(st["storage_compaction_seconds"] + st["lcp_finalize_seconds"], 3),
        "prefix_directory_k": info["prefix_directory_k"],
        "scope": ("SuffixArray::Build (HIP backend): text prepare, host->device through pinned staging, "
                  "device build, Fast prefix directory on the device, device->host, host storage"),
    }
    del index, ref
    return out


def timed_device_build(builder, d_text, n, bits, d_sa, d_isa, d_lcp, out_bits, flags, steps, warmup):
    def step():
        return builder.build(d_text, n, bits, d_sa, out_bits, d_isa, out_bits, d_lcp, out_bits, flags)

    for _ in range(warmup):
        step()
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(step())   # synchronous: returns after the build stream drains
    return stats, (time.perf_counter() - t0) / steps


def extra_leg(args, hip) -> dict:
    """The 64 Mbp multi-contig reference with 32-bit construction: device
    time, digests against the reference's, and the full check."""
    import oracle
    import sufkit_amd

    text, contigs = sufkit_amd.synthetic_text(2, args.seed, 0, 1.0)
    n, bases = int(text.size), int(contigs.sum())
    arrays = [sufkit_amd.DeviceArray(n + 64)] + [sufkit_amd.DeviceArray(n * 4) for _ in range(3)]
    arrays[0].upload(text)
    d_text, d_sa, d_isa, d_lcp = (a.ptr for a in arrays)
    with sufkit_amd.DeviceBuilder(0) as builder:
        flags = sufkit_amd.SFK_HIP_RETAIN_ISA | sufkit_amd.SFK_HIP_RETAIN_LCP
        stats, sec = timed_device_build(builder, d_text, n, 32, d_sa, d_isa, d_lcp, 32, flags,
                                        max(args.steps, 5), 1)
    out = {"workload": WORKLOADS[2], "construction_bits": 32, "text_symbols": n, "bases": bases,
           "ms_per_step": round(sec * 1e3, 3), "mbps": round(bases / 1e6 / sec, 2),
           "device_seconds": round(sum(s["device_seconds"] for s in stats) / len(stats), 6)}
    out.update(device_parity(d_text, n, d_sa, d_isa, d_lcp, 32, golden_digests(2, args.seed, 1.0),
                             f"0x{oracle.fingerprint(text[:-1]):016x}"))
    for a in arrays:
        a.free()
    return out


def main() -> int:
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); a text's build is "
              "not sharded across GPUs in this implementation", file=sys.stderr)
        return 2
    if args.workload == 4:
        return run_batched(args)
    if args.workload == 5:
        return run_served(args)
    import oracle
    import sufkit_amd

    text, contigs = sufkit_amd.synthetic_text(args.workload, args.seed, 0, args.scale)
    n = int(text.size)
    bases = int(contigs.sum())
    # The 3.1 Gbp genome is the 64-bit-construction workload (n > INT32_MAX,
    # the reference's divsufsort64 path); the others construct in 32 bits.
    bits = args.construction_bits or (64 if args.workload == 3 or n - 1 > 0xFFFFFFFF else 32)
    out_bits = 32 if n - 1 <= 0xFFFFFFFF else 64
    hip = sufkit_amd.load_hip()

    def dmalloc(nbytes):
        ptr = ctypes.c_void_p()
        if hip.sfk_hip_malloc(0, nbytes, ctypes.byref(ptr)) != 0:
            raise MemoryError(f"device allocation of {nbytes} bytes failed")
        return ptr.value

    d_text = dmalloc(n + 64)
    d_sa = dmalloc(n * out_bits // 8)
    d_isa = dmalloc(n * out_bits // 8)
    d_lcp = dmalloc(n * out_bits // 8)
    t0 = time.perf_counter()
    hip.sfk_hip_memcpy_h2d(d_text, text.ctypes.data, n)
    h2d = time.perf_counter() - t0

    if args.calibrate_copy:
        hip.sfk_hip_memcpy_d2d.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]
        size = args.calibrate_copy << 20
        src, dst = dmalloc(size), dmalloc(size)
        hip.sfk_hip_memcpy_d2d(dst, src, size)
        hip.sfk_hip_free(src)
        hip.sfk_hip_free(dst)
    tuning = radix_tuning(args, n)
    builder = sufkit_amd.DeviceBuilder(0)
    flags = sufkit_amd.SFK_HIP_RETAIN_ISA | sufkit_amd.SFK_HIP_RETAIN_LCP
    if not args.no_profile:
        flags |= sufkit_amd.SFK_HIP_PROFILE_KERNELS
    stats, sec_per_step = timed_device_build(builder, d_text, n, bits, d_sa, d_isa, d_lcp, out_bits, flags,
                                             args.steps, args.warmup)
    ms_per_step = sec_per_step * 1e3
    builder.close()  # its workspace is not needed by the checks and legs below

    # Results back to the host once (PCIe, outside the timed region), for
    # the element-by-element comparison with the CPU baselines.
    dt = np.uint32 if out_bits == 32 else np.uint64
    full_copy = n <= (1 << 28)
    sa = isa = lcp = None
    d2h = 0.0
    if full_copy:
        sa = np.empty(n, dt)
        isa = np.empty(n, dt)
        lcp = np.empty(n, dt)
        t0 = time.perf_counter()
        for arr, ptr in ((sa, d_sa), (isa, d_isa), (lcp, d_lcp)):
            hip.sfk_hip_memcpy_d2h(arr.ctypes.data, ptr, arr.nbytes)
        d2h = time.perf_counter() - t0
    fingerprint = f"0x{oracle.fingerprint(text[:-1]):016x}"
    parity = device_parity(d_text, n, d_sa, d_isa, d_lcp, out_bits,
                           golden_digests(args.workload, args.seed, args.scale), fingerprint)

    last = stats[-1]
    scatter_s = sum(s["sort_scatter_seconds"] for s in stats)
    scatter_n = sum(s["sort_scatter_launches"] for s in stats)
    scatter_b = sum(s["sort_scatter_bytes"] for s in stats)
    device_s = sum(s["device_seconds"] for s in stats) / len(stats)
    algo_bytes = last["algorithmic_bytes_sa"] + last["algorithmic_bytes_isa"] + last["algorithmic_bytes_lcp"]
    roofline = None
    if scatter_n:
        achieved = scatter_b / scatter_s / 1e9
        roofline = {
            "bound": "hbm", "kernel": "RadixOnesweep (LSD scatter pass, decoupled look-back)",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
            "launches_per_step": scatter_n // len(stats),
            "avg_launch_us": round(scatter_s / scatter_n * 1e6, 2),
            "bytes_per_launch": scatter_b // scatter_n,
        }
        if args.scale == 1.0:
            roofline.update(measured_traffic(args.workload, bits))

    result = {
        "metric": "SA+ISA+LCP build Mbp/s + HBM roofline frac, 1 MI355X vs divsufsort64 on host",
        "value": round(bases / 1e6 / sec_per_step, 2),
        "unit": "Mbp/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"u{bits}",
        "data": "synthetic (deterministic splitmix64 generator, seed %d)" % args.seed,
        "config": {
            "workload": WORKLOADS[args.workload] + (f" x{args.scale}" if args.scale != 1.0 else ""),
            "text_symbols": n, "bases": bases, "contigs": int(contigs.size),
            "construction_bits": bits, "stored_bits": out_bits,
            "profile": "fast (SA+ISA+raw LCP)", "fingerprint": fingerprint,
        },
        "roofline": roofline,
        "phases": {
            "device_seconds": round(device_s, 6),
            "sa_seconds": round(last["sa_seconds"], 6),
            "isa_seconds": round(last["isa_seconds"], 6),
            "lcp_seconds": round(last["lcp_seconds"], 6),
            "per_phase": {
                ph: {"seconds": round(last[f"{ph}_seconds"], 6),
                     "algorithmic_bytes": last[f"algorithmic_bytes_{ph}"],
                     "GBps": round(last[f"algorithmic_bytes_{ph}"] / max(last[f"{ph}_seconds"], 1e-12) / 1e9, 1),
                     "frac": round(last[f"algorithmic_bytes_{ph}"] / max(last[f"{ph}_seconds"], 1e-12) / 1e9
                                   / HBM_PEAK_GBPS, 4)}
                for ph in ("sa", "isa", "lcp")
            },
            "per_phase_note": ("isa: the ISA is the SA phase's rank array, written in place by its scatter "
                               "(those bytes count in sa); the isa phase holds only stored-width conversion "
                               "or the permutation check when there is no LCP pass"),
            "algorithmic_bytes": algo_bytes,
            "whole_build_GBps": round(algo_bytes / device_s / 1e9, 1),
            "whole_build_frac": round(algo_bytes / device_s / 1e9 / HBM_PEAK_GBPS, 4),
            "refine_steps": last["refine_steps"], "doubling_rounds": last["doubling_rounds"],
            "active_total": last["active_total"], "sort_passes": last["sort_passes"],
            "lce_long_items": last["lce_long_items"],
            "lcp_gathered_rows": last["lcp_gathered_rows"],
        },
        "parity": parity,
        "pcie_seconds": {"h2d_text": round(h2d, 4), "d2h_results": round(d2h, 4) if full_copy else None},
        "radix_tuning": tuning,
    }

    if not args.no_cpu_baseline:
        sample_symbols = args.cpu_sample_symbols or (0 if n <= (1 << 27) else 64_000_000)
        base, (rsa, risa, rlcp, whole, sample) = cpu_baseline(text, sample_symbols, bits)
        result["cpu_baseline"] = base
        if whole and full_copy:
            ok = (np.array_equal(sa.astype(np.uint64), rsa) and np.array_equal(isa.astype(np.uint64), risa)
                  and np.array_equal(lcp.astype(np.uint64), rlcp))
            result["parity_vs_cpu_baseline"] = "bit-exact" if ok else "MISMATCH"
        del rsa, risa, rlcp
        if oracle.ref_caps_available() and sample.size >= 16:
            caps, (csa, cisa, clcp) = cpu_baseline_caps(sample, bits, whole)
            result["cpu_baseline_caps"] = caps
            if whole and full_copy:
                ok = (np.array_equal(sa.astype(np.uint64), csa) and np.array_equal(isa.astype(np.uint64), cisa)
                      and np.array_equal(lcp.astype(np.uint64), clcp))
                result["parity_vs_cpu_baseline_caps"] = "bit-exact" if ok else "MISMATCH"
            del csa, cisa, clcp
    else:
        result["cpu_baseline"] = None

    if args.save:
        result["save"] = save_leg(args, text, contigs, d_text, d_sa, d_isa, d_lcp, bits, out_bits)
    if args.low_memory:
        result["low_memory"] = low_memory_leg(hip, d_sa, d_isa, d_lcp, n, out_bits, sa, lcp)
    for ptr in (d_text, d_sa, d_isa, d_lcp):
        hip.sfk_hip_free(ptr)
    del sa, isa, lcp
    if not args.no_host_build:
        result["host_build"] = host_build_leg(text, contigs, bits)
    if args.workload == 3 and args.scale == 1.0 and not args.no_extra_leg:
        result["extra_leg_64mbp"] = extra_leg(args, hip)
    print(json.dumps(result), flush=True)
    if args.save and result["save"].get("parity_vs_reference_save", "sections identical") != "sections identical":
        return 1
    if args.low_memory and result["low_memory"].get("decodes_to_raw_lcp", True) is not True:
        return 1
    legs = [parity] + ([result["extra_leg_64mbp"]] if "extra_leg_64mbp" in result else [])
    ok = (result.get("parity_vs_cpu_baseline", "bit-exact") == "bit-exact"
          and result.get("parity_vs_cpu_baseline_caps", "bit-exact") == "bit-exact"
          and all(leg.get("parity_vs_reference_digests", "bit-exact") == "bit-exact"
                  and not leg["full_check"]["failed"] for leg in legs))
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())

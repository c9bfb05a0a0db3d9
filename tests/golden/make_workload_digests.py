"""Write tests/golden/workload_digests.json (run in the build container).

Full-size golden digests of the benchmark workloads, from the reference's own
construction code compiled from its sources by oracle/Makefile:

  c1  4 Mbp i.i.d. genome            BuildDivsufsort32 (SA, ISA, Kasai LCP)
  c2  64 Mbp multi-contig genome     BuildDivsufsort32
  c4  512 bacterial-like references  BuildDivsufsort32, one per reference
  c3  3.1 Gbp human-scale genome     the reference's divsufsort64 SA (its
                                     64-bit construction), narrowed to 32
                                     bits; ISA and LCP by the oracle's ISA
                                     scatter and Phi/PLCP (the full 64-bit
                                     ISA/LCP arrays of BuildDivsufsort64 do
                                     not fit this host's memory)

Each entry holds the text's size and FNV-1a-64 fingerprint (so a run on
another machine proves it generated the same text) and the position-keyed
sum digests (oracle.sum_digest) of SA, ISA and LCP, the form the device
digest kernel computes.  Texts come from the deterministic generator in
sufkit_amd/host/synthetic.cpp (seed 20260822).

Usage: python tests/golden/make_workload_digests.py [--only c1,c2,c4,c3] [--jobs 6]
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from concurrent.futures import ProcessPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

OUT = HERE / "workload_digests.json"
SEED = 20260822


def _entry(text, sa, isa, lcp, method: str, seconds: float) -> dict:
    import oracle

    return {"n": int(text.size), "fingerprint": f"0x{oracle.fingerprint(text[:-1]):016x}",
            "sa": f"0x{oracle.sum_digest(sa):016x}", "isa": f"0x{oracle.sum_digest(isa):016x}",
            "lcp": f"0x{oracle.sum_digest(lcp):016x}", "lcp_max": int(lcp.max()),
            "method": method, "seconds": round(seconds, 2)}


def ref32_digests(workload: int, ref: int) -> dict:
    import oracle
    import sufkit_amd

    text, _ = sufkit_amd.synthetic_text(workload, SEED, ref, 1.0, threads=1)
    t0 = time.perf_counter()
    sa, isa, lcp = oracle.fast_build(text, "ref32")
    return _entry(text, sa, isa, lcp, "reference BuildDivsufsort32", time.perf_counter() - t0)


def c3_digests() -> dict:
    import oracle
    import sufkit_amd

    text, _ = sufkit_amd.synthetic_text(3, SEED, 0, 1.0)
    fp = oracle.fingerprint(text[:-1])
    t0 = time.perf_counter()
    sa32, ds_seconds = oracle.ref_divsufsort64_u32(text)
    d = oracle.lean_digests(text, sa32)
    return {"n": int(text.size), "fingerprint": f"0x{fp:016x}", "sa": f"0x{d['sa']:016x}",
            "isa": f"0x{d['isa']:016x}", "lcp": f"0x{d['lcp']:016x}", "lcp_max": d["lcp_max"],
            "sa_head": [int(x) for x in sa32[:8]],
            "method": "reference divsufsort64 SA; oracle ISA scatter and Phi/PLCP LCP",
            "divsufsort64_seconds": round(ds_seconds, 1),
            "seconds": round(time.perf_counter() - t0, 1)}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default="c1,c2,c4,c3")
    ap.add_argument("--jobs", type=int, default=6)
    ap.add_argument("--refs", type=int, default=512)
    args = ap.parse_args()
    import oracle

    if not oracle.ref_available():
        raise SystemExit("oracle/_ref is not built (needs the reference sources)")
    wanted = args.only.split(",")

    def save(key, value):
        # re-read before writing: several runs may fill different keys
        data = json.loads(OUT.read_text()) if OUT.exists() else {}
        data["seed"] = SEED
        data["digest"] = "sum digest: n*G + sum_i fmix64(v[i] + G*(i+1)) mod 2^64 (oracle.c)"
        data[key] = value
        OUT.write_text(json.dumps(data, indent=1) + "\n")
        print(key, json.dumps(value if key != "c4" else value["per_reference"][:2]), flush=True)

    if "c1" in wanted:
        save("c1", ref32_digests(1, 0))
    if "c2" in wanted:
        save("c2", ref32_digests(2, 0))
    if "c4" in wanted:
        with ProcessPoolExecutor(args.jobs) as pool:
            rows = list(pool.map(ref32_digests, [4] * args.refs, range(args.refs)))
        save("c4", {"references": args.refs, "per_reference": rows})
    if "c3" in wanted:
        save("c3", c3_digests())


if __name__ == "__main__":
    main()

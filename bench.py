#!/usr/bin/env python3
"""Benchmark: S values exhausted per second on the census S in [1, 1e10).

One step is one complete census of the range (every S, every representation
triple, every 4-set extended), exactly the work of `s4 1 10000000000` in the
reference (panjanek/pairwise-square-sums, exhaustive/s4.c).  The whole range
is the input; nothing is cached between steps.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
                    [--lo 1] [--hi 10000000000] [--no-cpu] [--no-latency]

Also reported:
  roofline      the extension kernel (the dominant one) against the integer
                lane-op peak of the chip, with the op model of the census:
                ops = 256*4sets + 48*ext_cand + 96*ext_sqrt for that kernel;
  roofline_job  the whole census: 4*S + 12*pairs + 40*triples + 256*4sets +
                48*ext_cand + 96*ext_sqrt over the wall time;
  cpu_baseline  the reference s4 binary (oracle/_ref/s4, built from the
                reference's source) on this host's cores, over a bounded sample
                of the same range;
  latency_ms    p50/p99 of range queries (lengths log-uniform 1e3..1e8, seed 7)
                answered one at a time.
"""
import argparse
import json
import math
import os
import random
import re
import subprocess
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "pairwise-square-sums_amd"))

PEAK_LANE_OPS = 256 * 64 * 2.4e9      # 256 CUs x 64 lanes x 2.4 GHz, 32-bit integer lane-ops/s
GOLDEN = os.path.join(REPO, "tests", "golden", "golden.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def op_model(c):
    ext = 256 * c["4-sets"] + 48 * c["ext_cand"] + 96 * c["ext_sqrt"]
    rq = 4 * c["S_done"] + 12 * c["pairs"] + 40 * c["triples"]
    return rq, ext


def cpu_baseline(lo, hi, threads):
    """Time the reference s4 (built from exhaustive/s4.c) on a bounded, spread sample."""
    s4 = os.path.join(REPO, "oracle", "_ref", "s4")
    kind = "reference"
    if not os.path.exists(s4):
        s4 = os.path.join(REPO, "oracle", "_bin", "ps4_oracle")
        kind = "port"
        threads = 1
    if not os.path.exists(s4):
        return None
    span = hi - lo
    nsub, width = 10, min(10**8, max(1, span // 100))
    if kind == "port":
        width = min(width, 2 * 10**6)
    done = 0
    t_total = 0.0
    with tempfile.TemporaryDirectory() as tmp:
        for k in range(nsub):
            a = lo + (span * k) // nsub + (span // nsub - width) // 2
            args = [s4, str(a), str(a + width)] + ([str(threads), "20"] if kind == "reference" else [])
            t0 = time.perf_counter()
            r = subprocess.run(args, cwd=tmp, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, timeout=600)
            t_total += time.perf_counter() - t0
            if r.returncode != 0:
                return None
            m = re.search(rb"S_done=(\d+)", r.stderr)
            done += int(m.group(1)) if m else width
    return {"value": done / t_total, "unit": "S/s", "cores": threads, "kind": kind,
            "sample": f"{nsub} subranges of {width} S spread evenly over [{lo},{hi}), "
                      f"{'s4 with ' + str(threads) + ' threads' if kind == 'reference' else 'single-thread port'}, "
                      f"{t_total:.1f} s"}


def latency(eng, hi, n=100, seed=7):
    rng = random.Random(seed)
    lat = []
    for _ in range(n):
        length = int(10 ** rng.uniform(3, 8))
        lo = rng.randrange(1, max(2, hi - length))
        t0 = time.perf_counter()
        eng.search(lo, lo + length)
        lat.append((time.perf_counter() - t0) * 1e3)
    lat.sort()
    return {"p50": lat[len(lat) // 2], "p99": lat[min(len(lat) - 1, int(0.99 * len(lat)))], "n": n,
            "lengths": "log-uniform 1e3..1e8", "lo": f"uniform in [1,{hi})"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--lo", type=int, default=1)
    ap.add_argument("--hi", type=int, default=10**10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-latency", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the census shards by S range "
              f"across GPUs (disjoint S intervals, one engine per device), but this bench measures one device.", file=sys.stderr)
        sys.exit(2)
    import ps4
    lo, hi = args.lo, args.hi
    eng = ps4.Engine(S_hi_max=hi)
    for w in range(args.warmup):
        log(f"warmup {w + 1}/{args.warmup}: census [{lo},{hi})")
        eng.search(lo, hi)
    eng.kernel_time(reset=True)
    times = []
    res = None
    for k in range(args.steps):
        t0 = time.perf_counter()
        res, lines = eng.search(lo, hi)        # returns after the device work and the result lines
        times.append(time.perf_counter() - t0)
        log(f"step {k + 1}/{args.steps}: {times[-1] * 1e3:.1f} ms")
    ms_rq, ms_ext, batches = eng.kernel_time()
    c = res["counters"]
    wall = sum(times)
    value = c["S_done"] * args.steps / wall
    launches = max(1, batches)
    rq_ops, ext_ops = op_model(c)
    ext_ms = ms_ext / launches
    ext_achieved = ext_ops * args.steps / (ms_ext * 1e-3) if ms_ext > 0 else 0.0
    exact = None
    if os.path.exists(GOLDEN):
        gold = json.load(open(GOLDEN))
        rec = next((r for r in gold["runs"] if (r["lo"], r["hi"]) == (lo, hi)), None)
        if rec:
            exact = (rec["counters"] == c and rec["hist5"] == res["hist5"] and len(lines) == rec["n_lines"])
    out = {
        "metric": "elements/sec exhausted to 1e10 (bit-exact vs s4.c); p50/p99 range-query latency",
        "value": value,
        "unit": "S/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": f"plain census, every S in [{lo},{hi})", "S_lo": lo, "S_hi": hi,
                   "batches_per_step": launches // args.steps},
        "bit_exact_vs_s4": exact,
        "counters": c,
        "roofline": {"bound": "valu", "kernel": "k_ext", "achieved": ext_achieved / 1e12, "peak": PEAK_LANE_OPS / 1e12,
                     "unit": "Tlane-op/s", "frac": ext_achieved / PEAK_LANE_OPS,
                     "avg_launch_ms": ext_ms, "ops_per_launch": ext_ops / max(1, launches // args.steps),
                     "traffic": None},
        "roofline_job": {"achieved": (rq_ops + ext_ops) * args.steps / wall / 1e12, "peak": PEAK_LANE_OPS / 1e12,
                         "unit": "Tlane-op/s", "frac": (rq_ops + ext_ops) * args.steps / wall / PEAK_LANE_OPS,
                         "k_rq_ms_per_step": ms_rq / args.steps, "k_ext_ms_per_step": ms_ext / args.steps},
    }
    if not args.no_latency:
        log("range-query latency")
        out["latency_ms"] = latency(eng, hi)
    eng.close()
    if not args.no_cpu:
        threads = min(16, len(os.sched_getaffinity(0)))
        log(f"cpu baseline on {threads} threads")
        out["cpu_baseline"] = cpu_baseline(lo, hi, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

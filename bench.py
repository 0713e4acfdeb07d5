#!/usr/bin/env python3
"""Benchmark: S values exhausted per second on the census S in [1, 1e10).

One step is one complete census of the range (every S, every representation
triple, every 4-set extended), exactly the work of `s4 1 10000000000` in the
reference (panjanek/pairwise-square-sums, exhaustive/s4.c).  The whole range
is the input; nothing is cached between steps.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
                    [--lo 1] [--hi 10000000000] [--no-cpu] [--no-latency] [--no-workloads]
    python bench.py --workload batched|band|serving|elements
    (any of them) [--tune FILE]  tuning file from ps4/autotune.py
                  [--graph]      every batch launched as one captured hipGraph

The default workload is the plain census above.  The others print their own
JSON line:
  batched  4096 subranges of [1,1e11], lengths log-uniform in [1e4,1e9]
           (splitmix64 seed 42): all in one ps4_search_batch against one
           ps4_search per subrange; per-subrange results must be identical.
  band     [9e11, 1e12) as 1000 subranges of 1e8 in one batched call; whole-
           band counters against the reference's own logs
           (exhaustive/logs/r12_9.err + boundary_9.err), and a seeded sample of
           subranges bit-exact against the reference s4 run on this host.
  elements every 4-set with largest element <= 1e7 and the square-triple
           count (the search by largest element), against the reference's
           --dump4 output; the reference s4 on one thread beside it.
  serving  an open-loop Poisson stream (200 req/s, seed 7; 70% range queries
           lo in [1,1e12], length log-uniform in [1e3,1e8]; 30% triple queries
           c <= 1e9, d <= 1e12) through the continuous-batching request loop;
           p50/p99 latency, throughput, and a sample of requests checked
           against the same requests run alone.

Also reported:
  roofline      the extension kernel (the dominant one) against the integer
                lane-op peak of the chip, with the op model of the census:
                ops = 256*4sets + 48*ext_cand + 96*ext_sqrt for that kernel;
  roofline_job  the whole census: 4*S + 12*pairs + 40*triples + 256*4sets +
                48*ext_cand + 96*ext_sqrt over the wall time;
  cpu_baseline  the reference s4 (built from the reference's source into
                oracle/_ref/) on every CPU the process may use (nproc), over
                three subranges spread over the range, median of three runs of
                about 30 s, with the CPU model, nproc and the -march used;
  latency_ms    p50/p99 of range queries (lengths log-uniform 1e3..1e8, seed 7)
                answered one at a time;
  batched       the batched workload's S/s and its ratio to one call per
                subrange (per-subrange results compared);
  serving       the serving workload (1000 requests): p50/p99 latency,
                throughput, and results alone against in the mix.
"""
import argparse
import hashlib
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
# HBM bytes per k_ext launch, measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on the
# [1,1e10) census (3 launches: their average), each counter scaled by what it reads for a 2 GiB device copy
# (profiles/census_1e10_counters.md); reported for that workload only
K_EXT_TRAFFIC = {(1, 10**10): 27.59e9}
GOLDEN = os.path.join(REPO, "tests", "golden", "golden.json")
TUNE = None   # tuning file (--tune): launch shape and batch plan of every engine opened here
GRAPH = False  # --graph: every batch launched as one captured hipGraph


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def op_model(c):
    ext = 256 * c["4-sets"] + 48 * c["ext_cand"] + 96 * c["ext_sqrt"]
    rq = 4 * c["S_done"] + 12 * c["pairs"] + 40 * c["triples"]
    return rq, ext


def host_cpu():
    """(model name, nproc = CPUs this process may run on, os.cpu_count(), flags)."""
    model, flags = None, set()
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, timeout=30).stdout.decode()
        m = re.search(r"^Model name:\s*(.+)$", out, re.M)
        model = m.group(1).strip() if m else None
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if model is None and line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                if line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                    break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0)), os.cpu_count(), flags


def cpu_baseline(lo, hi, target_s=30.0, runs=3, nsub=3):
    """The reference s4 (built from exhaustive/s4.c into oracle/_ref/) on every
    CPU this process may use (nproc, s4's own default), over nsub subranges
    spread evenly over [lo, hi), each a separate s4 process timed from outside:
    two calibration runs, then `runs` runs sized to target_s seconds or more each,
    reported as their median S/s.  The -march=x86-64-v3 build is used where the
    host has AVX2, BMI2 and FMA."""
    model, nproc, ncpu, flags = host_cpu()
    ref = os.path.join(REPO, "oracle", "_ref")
    s4, march = os.path.join(ref, "s4_v3"), "x86-64-v3"
    if not ({"avx2", "bmi2", "fma"} <= flags and os.path.exists(s4)):
        s4, march = os.path.join(ref, "s4"), "x86-64"
    if not os.path.exists(s4):
        return None
    threads = nproc
    span = hi - lo

    def run(width):
        done, wall, busy = 0, 0.0, 0.0
        with tempfile.TemporaryDirectory() as tmp:
            for k in range(nsub):
                a = lo + (span * k) // nsub + (span // nsub - width) // 2
                t0 = time.perf_counter()
                r = subprocess.run([s4, str(a), str(a + width), str(threads), "20"], cwd=tmp,
                                   stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, timeout=600)
                wall += time.perf_counter() - t0
                if r.returncode != 0:
                    return None
                err = r.stderr.decode()
                done += int(re.search(r"S_done=(\d+)", err).group(1))
                # the per-thread compute seconds (exhaustive/s4.c:496): s4's own wall
                # is rounded up by its one-second monitor poll (:473)
                busy += max(float(x) for x in re.findall(r"thread\s+\d+:.*?([\d.]+)s\s*$", err, re.M) or ["0"])
        return done, wall, busy
    # s4's monitor polls once a second (exhaustive/s4.c:473), so each process's
    # wall time is rounded up by up to a second and a short run understates the
    # rate: calibrate twice (the second run about 15 s), size the timed runs
    # from it with 25 % to spare, and if a timed run still ends below target_s,
    # size them again from the shortest and repeat them once.
    cal = run(min(10**8, span // nsub))
    if cal is None:
        return None
    width = int(min(span // nsub, max(10**7, cal[0] / cal[1] * 15.0 / nsub)))
    cal = run(width)
    if cal is None:
        return None
    width = int(min(span // nsub, max(10**7, cal[0] / cal[1] * target_s * 1.25 / nsub)))
    res = []
    for attempt in range(2):
        log(f"cpu baseline: s4 ({march}) on {threads} threads, {runs} runs of {nsub} x {width} S")
        res = [run(width) for _ in range(runs)]
        if any(x is None for x in res):
            return None
        shortest = min(w for _, w, _ in res)
        if shortest >= target_s or width >= span // nsub:
            break
        width = int(min(span // nsub, width * target_s * 1.1 / shortest))
    rates = sorted(d / w for d, w, _ in res)
    return {"value": rates[len(rates) // 2], "unit": "S/s", "cores": threads, "kind": "reference",
            "sample": f"{nsub} subranges of {width} S spread evenly over [{lo},{hi}), one s4 process each, "
                      f"median of {runs} runs ({', '.join(f'{w:.1f}' for _, w, _ in res)} s)",
            "runs_S_per_s": rates, "compute_S_per_s": sorted(d / b for d, _, b in res if b > 0),
            "binary": f"oracle/_ref/{os.path.basename(s4)}: gcc -O3 -march={march} exhaustive/s4.c",
            "cpu_model": model, "nproc": nproc, "cpu_count": ncpu}


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


def lines_sha(lines):
    return hashlib.sha256("".join(sorted(x + "\n" for x in lines)).encode()).hexdigest()


def s4_ref_run(lo, hi, threads):
    """The reference s4 (oracle/_ref/s4, built from exhaustive/s4.c) on [lo, hi):
    (counters dict, sorted-lines sha) or None when it is not built."""
    s4 = os.path.join(REPO, "oracle", "_ref", "s4")
    if not os.path.exists(s4):
        return None
    with tempfile.TemporaryDirectory() as tmp:
        outp = os.path.join(tmp, "o.out")
        r = subprocess.run([s4, str(lo), str(hi), str(threads), "20", "--out", outp], cwd=tmp,
                           stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, timeout=900)
        if r.returncode != 0:
            return None
        tot = re.search(rb"^TOTAL: (.*)$", r.stderr, re.M).group(1).decode()
        kv = {k: int(v) for k, v in re.findall(r"([\w-]+)=(\d+)", tot)}
        import ps4
        counters = {k: kv[k] for k in ps4.TOTAL_KEYS}
        lines = open(outp).read().splitlines()
    return counters, lines_sha(lines)


def copy_probe(mib):
    """One device-to-device hipMemcpy of mib MiB (the runtime's copy kernel),
    between two buffers filled by hipMemset."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_size_t(mib << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), n) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, n)
    hip.hipMemset(b, 0, n)
    hip.hipDeviceSynchronize()
    rc = hip.hipMemcpy(b, a, n, 3)   # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed ({rc})")
    log(f"copy probe: {mib} MiB device-to-device")


def add_counters(acc, c):
    for k, v in c.items():
        acc[k] = acc.get(k, 0) + v
    return acc


def bench_batched(args):
    import ps4
    from ps4 import workloads
    ranges = workloads.batched_ranges(4096, seed=42, top=10**11)
    los, his = [a for a, _ in ranges], [b for _, b in ranges]
    total = sum(b - a for a, b in ranges)
    eng = ps4.Engine(S_hi_max=10**11 + 1, tune=TUNE, graph=GRAPH)
    eng.search(1, 10**6)                     # warm the device
    log(f"batched: {len(ranges)} subranges, {total:.3e} S in one ps4_search_batch")
    eng.kernel_time(reset=True)
    t0 = time.perf_counter()
    res_b = eng.search_batch(los, his)
    t_b = time.perf_counter() - t0
    kb = eng.kernel_time(reset=True)
    log(f"batched: {t_b:.2f} s; plain: one ps4_search per subrange")
    t0 = time.perf_counter()
    res_p = [eng.search(a, b) for a, b in ranges]
    t_p = time.perf_counter() - t0
    kp = eng.kernel_time(reset=True)
    eng.close()
    same = all(x[0] == y[0] and sorted(x[1]) == sorted(y[1]) for x, y in zip(res_b, res_p))
    tot = {}
    for c, _ in res_b:
        add_counters(tot, c["counters"])
    return {
        "metric": "elements/sec, 4096 uneven subranges of [1,1e11] in one batched call vs one call per subrange",
        "value": total / t_b, "unit": "S/s", "n_gpus": 1, "higher_is_better": True, "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "batched: 4096 subranges [lo,lo+len) in [1,1e11], len = floor(10^(4+5U)), "
                               "lo = 1+floor(U'(1e11-len)), U,U' from splitmix64 seed 42",
                   "subranges": len(ranges), "S_total": total},
        "seconds_batched": t_b, "seconds_plain": t_p, "plain_value": total / t_p, "batched_over_plain": t_p / t_b,
        "per_subrange_identical": same, "counters": tot,
        "kernel_ms_batched": {"k_rq": kb[0], "k_ext": kb[1], "launch_pairs": kb[2]},
        "kernel_ms_plain": {"k_rq": kp[0], "k_ext": kp[1], "launch_pairs": kp[2]},
    }


def bench_band(args):
    import ps4
    from ps4 import workloads
    lo, hi, width = 9 * 10**11, 10**12, 10**8
    sub = workloads.band_ranges(lo, hi, width)
    eng = ps4.Engine(S_hi_max=hi, tune=TUNE, graph=GRAPH)
    eng.search(lo, lo + 10**6)
    eng.kernel_time(reset=True)
    log(f"band: [{lo},{hi}) as {len(sub)} subranges in one ps4_search_batch")
    t0 = time.perf_counter()
    res = eng.search_batch([a for a, _ in sub], [b for _, b in sub])
    t = time.perf_counter() - t0
    kt = eng.kernel_time(reset=True)
    eng.close()
    tot = {}
    n5_le = 0
    for c, lines in res:
        add_counters(tot, c["counters"])
        for ln in lines:   # the fifth element e <= 1e12 (members of a 5-set line, largest first)
            if ln.startswith("5:") and int(ln[4:].split(",", 1)[0]) <= 10**12:
                n5_le += 1
    out = {
        "metric": "elements/sec over the band [9e11,1e12) in batched form",
        "value": (hi - lo) / t, "unit": "S/s", "n_gpus": 1, "higher_is_better": True, "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "band [9e11,1e12) as 1000 subranges of 1e8, one batched call", "S_lo": lo, "S_hi": hi},
        "seconds": t, "counters": tot, "five_sets_with_largest_le_1e12": n5_le,
        "kernel_ms": {"k_rq": kt[0], "k_ext": kt[1], "launch_pairs": kt[2]},
    }
    if os.path.exists(GOLDEN):
        gold = json.load(open(GOLDEN))
        logs = {r["source"].rsplit("/", 1)[-1]: r for r in gold["reference_logs"]}
        if "r12_9.err" in logs and "boundary_9.err" in logs:
            ref = add_counters(dict(logs["r12_9.err"]["counters"]), logs["boundary_9.err"]["counters"])
            # the logs come from an earlier build of s4 than exhaustive/s4.c: the counters up
            # to the triples must agree exactly, the rest are reported beside them
            out["vs_reference_logs"] = {
                "exact_through_triples": all(ref[k] == tot[k] for k in ("S_done", "pairs", "S_ge3", "triples")),
                "log": ref, "difference": {k: tot[k] - ref[k] for k in ref}}
    if not args.no_cpu:
        rng = random.Random(20261203)
        pick = sorted(rng.sample(range(len(sub)), 3))
        threads = min(16, len(os.sched_getaffinity(0)))
        checks = []
        for i in pick:
            a, b = sub[i]
            log(f"band: s4 on [{a},{b}) with {threads} threads")
            t0 = time.perf_counter()
            r = s4_ref_run(a, b, threads)
            dt = time.perf_counter() - t0
            if r is None:
                checks = None
                break
            checks.append({"lo": a, "hi": b, "s4_seconds": dt, "identical": r[0] == res[i][0]["counters"]
                           and r[1] == lines_sha(res[i][1])})
        out["sample_vs_s4"] = checks
    return out


def bench_serving(args):
    import ps4
    from ps4 import serve, workloads
    n = args.requests
    stream = workloads.serving_stream(n, rate=200.0, seed=7)
    eng = ps4.Engine(S_hi_max=10**12 + 10**8 + 1, tune=TUNE, graph=GRAPH)
    eng.search(10**12 - 10**6, 10**12)   # warm the device
    srv = serve.Server(eng)
    log(f"serving: {n} requests at 200/s (open loop)")
    t_start = time.perf_counter()
    reqs = []
    for t_arr, kind, a in stream:
        now = time.perf_counter() - t_start
        if t_arr > now:
            time.sleep(t_arr - now)
        reqs.append(srv.submit(kind, a, t_arrive=t_start + t_arr))
    for r in reqs:
        r.wait()
    t_end = time.perf_counter()
    srv.close()
    errors = [repr(r.error) for r in reqs if r.error is not None]

    def pct(xs, q):
        xs = sorted(xs)
        return xs[min(len(xs) - 1, int(q * len(xs)))] * 1e3 if xs else None
    lat = [r.latency for r in reqs]
    lat_r = [r.latency for r in reqs if r.kind == "range"]
    lat_t = [r.latency for r in reqs if r.kind == "triple"]
    s_served = sum(r.args[1] - r.args[0] for r in reqs if r.kind == "range")
    # the same requests alone: a seeded sample re-run one by one
    rng = random.Random(11)
    sample = rng.sample(range(n), min(40, n))
    same = True
    for i in sample:
        r = reqs[i]
        if r.kind == "range":
            alone = eng.search(*r.args)
            same &= alone[0] == r.result[0] and sorted(alone[1]) == sorted(r.result[1])
        else:
            same &= eng.extend_triple(*r.args) == r.result
    eng.close()
    p50, p99 = pct(lat, 0.5), pct(lat, 0.99)
    return {
        "metric": "p50/p99 request latency at 200 req/s (range and triple queries), continuous batching",
        "value": p99, "unit": "ms (p99)", "n_gpus": 1, "higher_is_better": False, "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "serving: Poisson 200 req/s seed 7, 70% ranges lo~U[1,1e12] len~logU[1e3,1e8], "
                               "30% triples c<=1e9 d<=1e12", "requests": n},
        "latency_ms": {"p50": p50, "p99": p99, "p99_over_p50": p99 / p50 if p50 else None,
                       "range": {"p50": pct(lat_r, 0.5), "p99": pct(lat_r, 0.99)},
                       "triple": {"p50": pct(lat_t, 0.5), "p99": pct(lat_t, 0.99)}},
        "throughput": {"requests_per_s": n / (t_end - t_start), "range_S_per_s": s_served / (t_end - t_start)},
        "steps": len(srv.steps), "mean_requests_per_step": n / max(1, len(srv.steps)),
        "step_ms": {"p50": pct([t for _, t in srv.steps], 0.5), "p99": pct([t for _, t in srv.steps], 0.99)},
        "same_alone_and_in_mix": same, "sample_checked": len(sample), "errors": errors[:5],
    }


def bench_elements(args):
    """Element-range search (every 4-set with largest element <= N, and the
    square-triple count) at N = 1e7, against the reference fixture; the CPU leg
    is the reference s4 printing every 4-set of S < 4N - 5 on one thread
    (filtered by largest element afterwards)."""
    import ps4
    n_max = 10**7
    eng = ps4.Engine(S_hi_max=4 * n_max, tune=TUNE, graph=GRAPH)
    eng.search_elements(10**5)   # warm the device
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        lines, ntri = eng.search_elements(n_max)
        times.append(time.perf_counter() - t0)
    eng.close()
    t = min(times)
    out = {"metric": "element-range search: 4-sets with largest element <= 1e7 and square triples",
           "value": t * 1e3, "unit": "ms", "n_gpus": 1, "higher_is_better": False, "dtype": "u64",
           "data": "synthetic", "config": {"workload": "elements, N = 1e7", "n_max": n_max},
           "sets": len(lines), "square_triples": ntri, "times_s": times}
    if os.path.exists(GOLDEN):
        gold = json.load(open(GOLDEN))
        rec = next((r for r in gold.get("elements", []) if r["n_max"] == n_max), None)
        if rec:
            sha = hashlib.sha256("".join(x + "\n" for x in lines).encode()).hexdigest()
            out["bit_exact_vs_s4"] = len(lines) == rec["n_sets"] and sha == rec["sha256"]
    s4 = os.path.join(REPO, "oracle", "_ref", "s4")
    if not args.no_cpu and os.path.exists(s4):
        with tempfile.TemporaryDirectory() as tmp:
            t0 = time.perf_counter()
            r = subprocess.run([s4, "1", str(4 * n_max - 5), "1", "20", "--noext", "--dump4", str(2**62), "--out",
                                os.path.join(tmp, "o")], cwd=tmp, stdout=subprocess.DEVNULL,
                               stderr=subprocess.DEVNULL, timeout=600)
            dt = time.perf_counter() - t0
        if r.returncode == 0:
            out["cpu_baseline"] = {"value": dt * 1e3, "unit": "ms", "cores": 1, "kind": "reference",
                                   "sample": "the whole workload: s4 --noext --dump4 on [1, 4e7 - 5), one thread "
                                             "(square-triple count and the filter not included)"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--lo", type=int, default=1)
    ap.add_argument("--hi", type=int, default=10**10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-latency", action="store_true")
    ap.add_argument("--no-workloads", action="store_true",
                    help="plain census only (skip the batched and serving summaries)")
    ap.add_argument("--workload", choices=["plain", "batched", "band", "serving", "elements"], default="plain")
    ap.add_argument("--tune", default=None, help="tuning file from ps4/autotune.py (launch shape only)")
    ap.add_argument("--graph", action="store_true", help="launch every batch as one captured hipGraph")
    ap.add_argument("--requests", type=int, default=1000, help="serving: number of requests")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first copy this many MiB device-to-device once (a plain copy of known size that "
                         "calibrates the FETCH_SIZE / WRITE_SIZE counters in a profiled run)")
    args = ap.parse_args()
    global TUNE, GRAPH
    TUNE, GRAPH = args.tune, args.graph
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the census shards by S range "
              f"across GPUs (disjoint S intervals, one engine per device), but this bench measures one device.", file=sys.stderr)
        sys.exit(2)
    if args.workload != "plain":
        fn = {"batched": bench_batched, "band": bench_band, "serving": bench_serving,
              "elements": bench_elements}[args.workload]
        print(json.dumps(fn(args)), flush=True)
        return
    import ps4
    lo, hi = args.lo, args.hi
    eng = ps4.Engine(S_hi_max=hi, tune=TUNE, graph=GRAPH)
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib)
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
            exact = (rec["counters"] == c and rec["hist5"] == res["hist5"] and len(lines) == rec["n_lines"]
                     and lines_sha(lines) == rec["sorted_sha256"])
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
                     "traffic": K_EXT_TRAFFIC.get((lo, hi)), "traffic_unit": "B/launch",
                     "algorithmic_bytes_per_launch": 32 * c["4-sets"] / max(1, launches // args.steps)},
        "roofline_job": {"achieved": (rq_ops + ext_ops) * args.steps / wall / 1e12, "peak": PEAK_LANE_OPS / 1e12,
                         "unit": "Tlane-op/s", "frac": (rq_ops + ext_ops) * args.steps / wall / PEAK_LANE_OPS,
                         "k_rq_ms_per_step": ms_rq / args.steps, "k_ext_ms_per_step": ms_ext / args.steps},
    }
    if not args.no_latency:
        log("range-query latency")
        out["latency_ms"] = latency(eng, hi)
    eng.close()
    if not args.no_workloads:
        # the batched and serving workloads, summarised (each also has its own
        # --workload line with every detail)
        b = bench_batched(args)
        out["batched"] = {k: b[k] for k in ("value", "unit", "plain_value", "batched_over_plain",
                                             "per_subrange_identical", "seconds_batched", "seconds_plain")}
        out["batched"]["workload"] = b["config"]["workload"]
        sv = bench_serving(args)
        out["serving"] = {"latency_ms": sv["latency_ms"], "throughput": sv["throughput"],
                          "same_alone_and_in_mix": sv["same_alone_and_in_mix"], "requests": sv["config"]["requests"],
                          "workload": sv["config"]["workload"], "errors": sv["errors"]}
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(lo, hi)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

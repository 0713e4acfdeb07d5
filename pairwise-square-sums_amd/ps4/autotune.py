"""Autotuner of the census engine's launch shape and batch plan.

It varies the fields of ps4_tuning (include/ps4.h):
  - the launch shape: k_rq's block size (threads per workgroup: 128, 256,
    512) and per-thread tile (window points each thread loads at once: 2, 4,
    8), and k_ext's work-claim size (1, 4 or 16 tiles of 64 4-sets) -- each a
    compiled instantiation of the kernel;
  - the resident workgroups of k_rq and k_ext (per CU);
  - the subrange split (superwindow shift);
  - the batch fill of the 4-set queue.

The search is coordinate descent on a sample of the uneven-subrange workload:
n subranges of [1, top] with log-uniform lengths (ps4.workloads.batched_ranges).
Each setting is timed by wall clock over one ps4_search_batch call, best of
`reps`.  A setting is kept only if it is faster by more than `min_gain` and
every subrange's counters and result lines equal those of the default setting.
A tuning that changed a result would be a bug, so the tuner raises on one.

The result is a tuning file in the grammar ps4_parse_tuning reads, which
Engine(tune=path) and `ps4 --tune FILE` apply.

    PYTHONPATH=pairwise-square-sums_amd python -m ps4.autotune --out tune.txt
"""
import argparse
import json
import sys
import time

from . import Engine, format_tuning
from .workloads import batched_ranges


def _run(eng, los, his):
    t0 = time.perf_counter()
    res = eng.search_batch(los, his)
    return time.perf_counter() - t0, res


def candidates(key, cur, mx):
    """Values of one field tried around the one in force."""
    n_cu = max(1, mx.get("n_cu", 1))
    if key in ("rq_grid", "ext_grid"):
        per_cu = max(1, mx[key] // n_cu)
        return sorted({n_cu * b for b in range(max(1, per_cu - 3), per_cu + 1)} | {cur[key]})
    if key == "rq_block":
        return [128, 256, 512]
    if key == "rq_tile":
        return [2, 4, 8]
    if key == "ext_chunk":
        return [1, 4, 16]
    if key == "sw_shift":
        return list(range(0, min(3, mx[key]) + 1))
    if key == "batch_pct":
        return [40, 60, 80, 95]
    raise KeyError(key)


# the order of the descent: launch shape first (the grids' maxima depend on it)
KEYS = ("rq_block", "rq_tile", "ext_chunk", "ext_grid", "rq_grid", "sw_shift", "batch_pct")


def tune(eng, n=256, seed=42, top=10**11, e0=4, e1=9, reps=2, min_gain=0.01, rounds=1, log=print):
    """Coordinate descent over the tuning fields; returns (best tuning dict, report dict)."""
    ranges = batched_ranges(n, seed=seed, top=top, e0=e0, e1=e1)
    los, his = [a for a, _ in ranges], [b for _, b in ranges]
    eng.set_tuning({})                       # defaults
    cur, mx = eng.get_tuning()
    _run(eng, los[:8], his[:8])              # warm the device
    t_ref, ref = min((_run(eng, los, his) for _ in range(reps)), key=lambda x: x[0])
    log(f"autotune: {n} subranges, {sum(b - a for a, b in ranges):.3e} S; default {t_ref:.3f} s {cur}")
    best = {k: cur[k] for k in KEYS}
    t_best = t_ref
    trials = []
    for _ in range(rounds):
        changed = False
        for key in KEYS:
            for v in candidates(key, best, mx):
                if v == best[key]:
                    continue
                trial = dict(best, **{key: v})
                if key == "rq_block" or key == "rq_tile":
                    trial["rq_grid"] = 0          # the resident grid of the new shape
                eng.set_tuning(trial)
                trial = {k: eng.get_tuning()[0][k] for k in KEYS}   # (0 fields resolved)
                t, res = min((_run(eng, los, his) for _ in range(reps)), key=lambda x: x[0])
                if res != ref:
                    bad = next(i for i in range(n) if res[i] != ref[i])
                    raise RuntimeError(f"tuning {trial} changed the result of subrange {ranges[bad]}")
                trials.append({"tuning": trial, "seconds": t})
                log(f"autotune: {key}={v}: {t:.3f} s (best {t_best:.3f} s)")
                if t < t_best * (1.0 - min_gain):
                    best, t_best, changed = trial, t, True
        if not changed:
            break
    eng.set_tuning(best)
    report = {"sample": {"n": n, "seed": seed, "top": top, "len_log10": [e0, e1]}, "n_cu": mx.get("n_cu"),
              "default_seconds": t_ref, "best_seconds": t_best, "best": best, "trials": trials}
    return best, report


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", required=True, help="tuning file to write")
    ap.add_argument("--report", help="JSON report of every trial")
    ap.add_argument("--sample", type=int, default=256, help="subranges in the timed sample")
    ap.add_argument("--top", type=float, default=1e11, help="subranges lie in [1, top]")
    ap.add_argument("--e1", type=float, default=9.0, help="longest subrange 10^e1")
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--rounds", type=int, default=1)
    ap.add_argument("--device", type=int, default=0)
    args = ap.parse_args(argv)
    top = int(args.top)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)
    with Engine(S_hi_max=top + 1, device=args.device) as eng:
        best, report = tune(eng, n=args.sample, top=top, e1=args.e1, reps=args.reps, rounds=args.rounds, log=log)
    note = (f"ps4 autotune: {args.sample} subranges of [1,{top}] (seed 42), n_cu {report['n_cu']}\n"
            f"default {report['default_seconds']:.4f} s, tuned {report['best_seconds']:.4f} s")
    with open(args.out, "w") as f:
        f.write(format_tuning(best, note))
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)
    print(json.dumps({"best": best, "default_seconds": report["default_seconds"],
                      "best_seconds": report["best_seconds"]}))
    return 0


if __name__ == "__main__":
    sys.exit(main())

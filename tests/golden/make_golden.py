#!/usr/bin/env python3
"""Write the golden fixtures under tests/golden/ from the reference s4 binary.

The binary is oracle/_ref/s4, compiled by `make -C oracle ref` from the
reference's single source file (exhaustive/s4.c).  This script only runs it on
fixed ranges and records what it prints: the TOTAL counter vector
(exhaustive/s4.c:498-500), the 5-set histogram (:501-503), the CAPACITY maxima
(:504-506), the number of result lines, the SHA-256 of the byte-sorted lines
and, when the output is small, the sorted lines themselves.

It also records counter vectors that the reference ships as data
(exhaustive/logs/*.err) and the known-answer lines of exhaustive/test_verify.py
and bruteforce-scratchpad/dotnet/1mln.txt.

Usage:  python3 tests/golden/make_golden.py [--big]
        --big adds [1,1e9) and [1,1e10) (about 4 minutes on 8 cores).
        python3 tests/golden/make_golden.py --elements N
        runs the reference once more on [1, 4N-5) with every 4-set printed and
        records the 4-sets with largest element <= N (the element-range form
        of the search), their count and the SHA-256 of their "[a,b,c,d]"
        lines, members and list ascending;
        python3 tests/golden/make_golden.py --square-triples
        adds to each element record the number of square triples with
        c <= N, enumerated by the oracle (seconds);
        python3 tests/golden/make_golden.py --add-pooled LO HI PIECES WORKERS
        runs [LO, HI) as PIECES single-thread pieces on WORKERS processes and
        adds the summed record (the long ranges: [1, 1e11) this way);
        python3 tests/golden/make_golden.py --batched-sample WORKERS
        runs a seeded sample of 32 of the batched workload's subranges;
        python3 tests/golden/make_golden.py --band-sample WORKERS
        runs a seeded sample of 3 of the band's 1e8-wide subranges;
        python3 tests/golden/make_golden.py --add LO HI
        runs one more range (8 threads) and adds it to the existing fixture;
        [1, 6250000000) is there this way (about 2.5 minutes on 8 cores): it is
        the first thread's range of exhaustive/logs/s4d_1e11.err, whose
        4-set and extension counters come from an earlier build of the program
        than exhaustive/s4.c (its TOTAL line carries "nonprimitive skipped=",
        which s4.c does not print).
"""
import hashlib
import json
import os
import random
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
S4 = os.path.join(REPO, "oracle", "_ref", "s4")
REF = "/root/reference"
OUT = os.path.join(HERE, "golden.json")

KEYS = ["S_done", "pairs", "S_ge3", "triples", "4-sets", "5-sets", "6-sets",
        "multi-ext-4sets", "ext_cand", "ext_sqrt"]
MAX_INLINE_LINES = 2000


def parse_err(text):
    tot = re.search(r"^TOTAL: (.*)$", text, re.M).group(1)
    kv = {k: int(v) for k, v in re.findall(r"([\w-]+)=(\d+)", tot)}
    counters = {k: kv[k] for k in KEYS}
    hist = {}
    m = re.search(r"^5-sets by log2\(largest element\)[^:]*:(.*)$", text, re.M)
    if m:
        for a, b in re.findall(r"(\d+):(\d+)", m.group(1)):
            hist[a] = int(b)
    cap = re.search(r"^CAPACITY: reps=(\d+)/\d+ divisors=(\d+)/\d+ extensions=(\d+)/\d+", text, re.M)
    capacity = None
    if cap:
        capacity = {"reps": int(cap.group(1)), "divisors": int(cap.group(2)), "extensions": int(cap.group(3))}
    return counters, hist, capacity


def run_s4(lo, hi, threads, extra=()):
    with tempfile.TemporaryDirectory() as tmp:
        outp = os.path.join(tmp, "o.out")
        cmd = [S4, str(lo), str(hi), str(threads), "20", "--out", outp] + list(extra)
        res = subprocess.run(cmd, cwd=tmp, stdout=subprocess.PIPE, stderr=subprocess.PIPE, check=True)
        with open(outp, "rb") as f:
            lines = f.read().splitlines(keepends=True)
    lines.sort()
    counters, hist, capacity = parse_err(res.stderr.decode())
    rec = {"lo": lo, "hi": hi, "flags": list(extra), "threads": threads, "counters": counters,
           "hist5": hist, "capacity": capacity, "n_lines": len(lines),
           "sorted_sha256": hashlib.sha256(b"".join(lines)).hexdigest()}
    if len(lines) <= MAX_INLINE_LINES:
        rec["lines"] = [ln.decode().rstrip("\n") for ln in lines]
    return rec, lines


def reference_logs():
    """Counter vectors held by the reference in exhaustive/logs/*.err."""
    logs = os.path.join(REF, "exhaustive", "logs")
    recs = []
    for name in sorted(os.listdir(logs)):
        if not name.endswith(".err"):
            continue
        text = open(os.path.join(logs, name)).read()
        hdr = re.search(r"s4: S in \[(\d+),\s*(\d+)\)", text)
        counters, hist, capacity = parse_err(text)
        recs.append({"source": "exhaustive/logs/" + name, "lo": int(hdr.group(1)), "hi": int(hdr.group(2)),
                     "counters": counters, "hist5": hist})
    return recs


def add_run(lo, hi):
    data = json.load(open(OUT))
    rec, _ = run_s4(lo, hi, 8)
    data["runs"] = [r for r in data["runs"] if (r["lo"], r["hi"]) != (lo, hi)] + [rec]
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    print(lo, hi, rec["counters"]["4-sets"], rec["n_lines"], "added to", OUT)


def _piece(args):
    lo, hi = args
    cache = os.environ.get("GOLDEN_PIECE_CACHE")   # a directory: finished pieces survive a restart
    if cache:
        path = os.path.join(cache, f"{lo}_{hi}.json")
        if os.path.exists(path):
            with open(path) as f:
                d = json.load(f)
            return lo, hi, d["err"], [x.encode() for x in d["lines"]]
    with tempfile.TemporaryDirectory() as tmp:
        outp = os.path.join(tmp, "o.out")
        res = subprocess.run([S4, str(lo), str(hi), "1", "20", "--out", outp], cwd=tmp,
                             stdout=subprocess.PIPE, stderr=subprocess.PIPE, check=True)
        with open(outp, "rb") as f:
            lines = f.read().splitlines(keepends=True)
    if cache:
        with open(path + ".tmp", "w") as f:
            json.dump({"err": res.stderr.decode(), "lines": [x.decode() for x in lines]}, f)
        os.replace(path + ".tmp", path)
    return lo, hi, res.stderr.decode(), lines


def pooled_record(lo, hi, pieces, workers):
    """One range run as `pieces` equal single-thread s4 runs on a worker pool.
    The reference's counters are additive over a partition of the range and
    its capacity figures are maxima (exhaustive/s4.c:488-506), so the sums,
    maxima and the concatenated, byte-sorted lines equal one run's."""
    from multiprocessing import Pool
    cuts = [lo + (hi - lo) * k // pieces for k in range(pieces + 1)]
    jobs = [(cuts[k], cuts[k + 1]) for k in range(pieces) if cuts[k] < cuts[k + 1]]
    jobs.sort(key=lambda j: -j[0])            # costly high-S pieces first
    counters = {k: 0 for k in KEYS}
    hist, capacity, lines = {}, {"reps": 0, "divisors": 0, "extensions": 0}, []
    with Pool(workers) as pool:
        for n, (a, b, err, ls) in enumerate(pool.imap_unordered(_piece, jobs), 1):
            c, h, cap = parse_err(err)
            for k in KEYS:
                counters[k] += c[k]
            for k, v in h.items():
                hist[k] = hist.get(k, 0) + v
            for k in capacity:
                capacity[k] = max(capacity[k], cap[k])
            lines += ls
            print("piece %d/%d [%d,%d) done" % (n, len(jobs), a, b), flush=True)
    lines.sort()
    rec = {"lo": lo, "hi": hi, "flags": [], "threads": 1, "pieces": len(jobs), "counters": counters,
           "hist5": {k: hist[k] for k in sorted(hist, key=int)}, "capacity": capacity,
           "n_lines": len(lines), "sorted_sha256": hashlib.sha256(b"".join(lines)).hexdigest()}
    if len(lines) <= MAX_INLINE_LINES:
        rec["lines"] = [ln.decode().rstrip("\n") for ln in lines]
    return rec


def add_pooled(lo, hi, pieces, workers):
    rec = pooled_record(lo, hi, pieces, workers)
    data = json.load(open(OUT))                # re-read: other additions may have landed meanwhile
    data["runs"] = [r for r in data["runs"] if (r["lo"], r["hi"]) != (lo, hi)] + [rec]
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    print(lo, hi, rec["counters"], rec["n_lines"], "added to", OUT, flush=True)


def element_record(n_max):
    """The 4-sets with largest element <= n_max, from the reference's dump of
    every 4-set with S < 4 n_max - 5 (a + b + c + d <= 4 n_max - 6)."""
    _, lines = run_s4(1, 4 * n_max - 5, 1, ["--noext", "--dump4", "9223372036854775807"])
    sets = []
    for ln in lines:
        vals = sorted(int(x) for x in re.match(rb"4: \[([\d,]+)\]", ln).group(1).split(b","))
        if vals[3] <= n_max:
            sets.append(vals)
    sets.sort()
    text = "".join("[%d,%d,%d,%d]\n" % tuple(v) for v in sets)
    return {"n_max": n_max, "n_sets": len(sets), "sha256": hashlib.sha256(text.encode()).hexdigest(),
            "square_triples": enumerated_square_triples(n_max)}


def enumerated_square_triples(n_max):
    """Square triples {a<b<c} with c <= n_max, by direct enumeration of their
    root triples in the oracle (oracle/cpu_ref.c, ora_square_triples); the
    reference has no such count, so this is the fixture's only source."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import ora
    return ora.square_triples(n_max)


def add_batched_sample(workers):
    """s4 on a seeded sample of 32 of the 4096 uneven subranges of [1, 1e11]
    (the batched workload: splitmix64 seed 42, lengths log-uniform in
    [1e4, 1e9]); the sample indices come from random.Random(4096)."""
    sys.path.insert(0, os.path.join(REPO, "pairwise-square-sums_amd"))
    from ps4 import workloads
    ranges = workloads.batched_ranges(4096, seed=42, top=10**11)
    pick = sorted(random.Random(4096).sample(range(len(ranges)), 32))
    recs = []
    for i in pick:
        lo, hi = ranges[i]
        rec = pooled_record(lo, hi, max(1, (hi - lo) // (5 * 10**7)), workers)
        rec["index"] = i
        recs.append(rec)
        print("batched sample", i, lo, hi, rec["n_lines"], flush=True)
    data = json.load(open(OUT))
    data["batched_sample"] = {"workload": "batched_ranges(4096, seed=42, top=10**11)", "sample_seed": 4096,
                              "runs": recs}
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)


def add_band_sample(workers):
    """s4 on a seeded 0.3% sample (3 of 1000) of the 1e8-wide subranges of the
    band [9e11, 1e12); the sample indices come from random.Random(1000)."""
    lo, hi, width = 9 * 10**11, 10**12, 10**8
    pick = sorted(random.Random(1000).sample(range((hi - lo) // width), 3))
    recs = []
    for i in pick:
        a = lo + i * width
        rec = pooled_record(a, a + width, 8, workers)
        rec["index"] = i
        recs.append(rec)
        print("band sample", i, a, rec["n_lines"], flush=True)
    data = json.load(open(OUT))
    data["band_sample"] = {"band": [lo, hi], "width": width, "sample_seed": 1000, "runs": recs}
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)


def add_square_triple_counts():
    data = json.load(open(OUT))
    for rec in data.get("elements", []):
        rec["square_triples"] = enumerated_square_triples(rec["n_max"])
        print("elements", rec["n_max"], "square_triples", rec["square_triples"])
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)


def add_elements(n_max):
    data = json.load(open(OUT))
    rec = element_record(n_max)
    data["elements"] = [r for r in data.get("elements", []) if r["n_max"] != n_max] + [rec]
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    print("elements", rec, "added to", OUT)


def main():
    if "--band-sample" in sys.argv:
        return add_band_sample(int(sys.argv[sys.argv.index("--band-sample") + 1]))
    if "--batched-sample" in sys.argv:
        return add_batched_sample(int(sys.argv[sys.argv.index("--batched-sample") + 1]))
    if "--square-triples" in sys.argv:
        return add_square_triple_counts()
    if "--elements" in sys.argv:
        return add_elements(int(sys.argv[sys.argv.index("--elements") + 1]))
    if "--add" in sys.argv:
        i = sys.argv.index("--add")
        return add_run(int(sys.argv[i + 1]), int(sys.argv[i + 2]))
    if "--add-pooled" in sys.argv:
        i = sys.argv.index("--add-pooled")
        return add_pooled(int(sys.argv[i + 1]), int(sys.argv[i + 2]), int(sys.argv[i + 3]),
                          int(sys.argv[i + 4]))
    big = "--big" in sys.argv
    if not os.path.exists(S4):
        sys.exit("build the reference first: make -C oracle ref")
    data = {"generator": "tests/golden/make_golden.py", "runs": [], "dump4": [], "reference_logs": reference_logs()}
    plain = [(1, 10**5), (1, 4 * 10**6), (1, 10**7), (1, 10**8),
             (9 * 10**11, 9 * 10**11 + 10**7),
             (10**12, 10**12 + 2 * 10**6),
             (1954600000000, 1954600000000 + 10**6),
             (11762800000000, 11762800000000 + 3 * 10**5)]
    rng = random.Random(20261203)
    for _ in range(12):                       # scattered windows up to 1e12
        lo = rng.randrange(1, 10**12)
        plain.append((lo, lo + rng.choice([1, 7, 4096, 65536, 200003])))
    for lo in (2, 3, 5, 6, 25, 50, 65, 325, 1105, 4205):   # tiny ranges around small S with reps
        plain.append((lo, lo + 1))
    # the S of the known lines in exhaustive/test_verify.py:13 and :31-32
    plain += [(362141, 362142), (50578642000, 50578642001)]
    plain.append((1, 2))
    if big:
        plain += [(1, 10**9), (1, 10**10)]
    for lo, hi in plain:
        threads = 8 if hi - lo > 10**8 else 1
        rec, _ = run_s4(lo, hi, threads)
        data["runs"].append(rec)
        print(lo, hi, rec["counters"]["4-sets"], rec["n_lines"], flush=True)
    # --noext --dump4 runs: every 4-set, as printed by exhaustive/s4.c:237-241
    for lo, hi, inline in [(1, 10**5, True), (1, 4 * 10**7, False), (9 * 10**11, 9 * 10**11 + 2 * 10**4, True)]:
        rec, lines = run_s4(lo, hi, 1, ["--noext", "--dump4", "9223372036854775807"])
        if not inline:
            rec.pop("lines", None)
        # members of each 4-set are printed largest first; count the sets with largest element <= 1e7
        small = 0
        for ln in lines:
            vals = [int(x) for x in re.match(rb"4: \[([\d,]+)\]", ln).group(1).split(b",")]
            small += max(vals) <= 10**7
        rec["n_max_le_1e7"] = small
        data["dump4"].append(rec)
        print("dump4", lo, hi, rec["n_lines"], flush=True)
    data["known_lines"] = {
        "exhaustive/test_verify.py:13": "5: [763442,177458,148583,28658,7442] S=362141",
        "exhaustive/test_verify.py:31-32": "M: [28860026400,17124487200,4431625200,162503200] "
                                           "ext=[4981255200,2553991200,1674390681] S=50578642000",
    }
    onemln = os.path.join(REF, "bruteforce-scratchpad", "dotnet", "1mln.txt")
    if os.path.exists(onemln):
        # the only 5-sets with largest element below 10^6 (a JSON-like list with trailing commas)
        data["known_5sets_below_1e6"] = [[int(x) for x in m.split(",")]
                                         for m in re.findall(r"\[(\d+(?:,\d+)+)\]", open(onemln).read())]
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()

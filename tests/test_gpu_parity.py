"""Parity of the HIP census with the reference -- needs an MI355X (-m gpu).

Every test calls the engine through the C ABI (ctypes over libps4.so) or the
drop-in CLI built on it, and compares with
  * fixtures written by the reference s4.c (tests/golden/golden.json),
  * the counter vectors the reference ships (exhaustive/logs/*.err),
  * the CPU restatement (oracle/) on seeded windows the fixtures do not cover.
Bit-exact throughout: counters, 5-set histogram, capacity maxima, and the
byte-sorted result lines.  Tests run from small to large ranges.
"""
import hashlib
import math
import os
import random
import re
import subprocess
import sys
import time

import pytest

import ora
import ps4
from conftest import sha_sorted

pytestmark = pytest.mark.gpu

S_HI_MAX = 2 * 10**13


def say(msg):
    sys.__stderr__.write(f"[gpu-test] {msg}\n")
    sys.__stderr__.flush()


@pytest.fixture(scope="module")
def eng():
    say("open engine S_hi_max=2e13")
    e = ps4.Engine(S_hi_max=S_HI_MAX)
    yield e
    e.close()


def check(rec, res, lines, what=""):
    assert res["counters"] == rec["counters"], what
    assert res["hist5"] == rec["hist5"], what
    if rec.get("capacity"):
        assert res["capacity"] == rec["capacity"], what
    assert len(lines) == rec["n_lines"], what
    assert sha_sorted(lines) == rec["sorted_sha256"], what


def test_tiny_ranges(eng, golden):
    """Single S values and tiny windows (including S = 1..7 edge cases)."""
    for rec in golden["runs"]:
        if rec["hi"] - rec["lo"] <= 8:
            say(f"tiny [{rec['lo']},{rec['hi']})")
            res, lines = eng.search(rec["lo"], rec["hi"])
            check(rec, res, lines, (rec["lo"], rec["hi"]))


def test_single_s_4205_regression(eng, golden):
    """[4205, 4206): the first S with a 4-set ({3362, 482, 359, 2}).  During
    bring-up the GPU took a memory fault on it: residues mod 11..19 computed as
    __umul24(s, M) >> k were compiled to a 24-bit multiply followed by an
    arithmetic shift, so a product of 2^31 or more gave a negative residue that
    indexed far past a residue table.  Fixed with plain 32-bit multiplies; the
    residue inputs are clamped to the domains on which every residue is checked
    exhaustively (test_abi.py), so no index can leave its table, and test_abi.py
    checks the compiled kernels for arithmetic shifts."""
    rec = next(r for r in golden["runs"] if (r["lo"], r["hi"]) == (4205, 4206))
    say("single S [4205,4206)")
    res, lines = eng.search(4205, 4206)
    check(rec, res, lines)
    assert res["counters"]["4-sets"] == 1


def test_empty_range(eng):
    res, lines = eng.search(10**6, 10**6)
    assert res["counters"]["S_done"] == 0 and lines == []


def test_small_fixtures(eng, golden):
    for rec in golden["runs"]:
        if 8 < rec["hi"] - rec["lo"] <= 10**7 and rec["lo"] < S_HI_MAX:
            say(f"fixture [{rec['lo']},{rec['hi']})")
            res, lines = eng.search(rec["lo"], rec["hi"])
            check(rec, res, lines, (rec["lo"], rec["hi"]))
            if "lines" in rec:
                assert sorted(lines) == sorted(rec["lines"])


def test_readme_known_answer(eng):
    say("[1,1e7) known answer")
    res, lines = eng.search(1, 10**7)
    c = res["counters"]
    assert (c["5-sets"], c["multi-ext-4sets"], c["6-sets"]) == (123, 2, 0)


def test_reference_log_singletons(eng, golden):
    for r in golden["reference_logs"]:
        if r["source"].startswith("exhaustive/logs/boundary_"):
            say(r["source"])
            res, _ = eng.search(r["lo"], r["hi"])
            assert res["counters"] == r["counters"], r["source"]


def test_seeded_windows_vs_oracle(eng):
    """Windows the fixtures do not cover, against the CPU restatement."""
    rng = random.Random(42)
    for i in range(10):
        mag = rng.choice([10**6, 10**8, 10**10, 10**11, 10**12])
        lo = rng.randrange(1, mag)
        span = rng.choice([1, 17, 4096, 16383, 16384, 16385, 65537, 150000])
        say(f"oracle window {i}: [{lo},{lo + span})")
        res, lines = eng.search(lo, lo + span)
        want, wl = ora.census(lo, lo + span)
        assert res == want, (lo, span)
        assert sorted(lines) == wl, (lo, span)


def test_dump4(golden):
    e = ps4.Engine(S_hi_max=10**12, noext=True, dump4=2**62)
    try:
        for rec in golden["dump4"]:
            say(f"dump4 [{rec['lo']},{rec['hi']})")
            res, lines = e.search(rec["lo"], rec["hi"])
            assert res["counters"] == rec["counters"]
            assert len(lines) == rec["n_lines"]
            assert sha_sorted(lines) == rec["sorted_sha256"]
            if "n_max_le_1e7" in rec:
                small = sum(max(int(x) for x in ln.split("[")[1].split("]")[0].split(",")) <= 10**7 for ln in lines)
                assert small == rec["n_max_le_1e7"]
    finally:
        e.close()


def test_dump4_limit_takes_first_in_s_order():
    """--dump4 N prints the first N 4-sets in the reference's single-thread order."""
    e = ps4.Engine(S_hi_max=10**7, noext=True, dump4=50)
    try:
        say("dump4 limit 50")
        _, lines = e.search(1, 10**6)
    finally:
        e.close()
    _, first = ora.census(1, 10**6, noext=True, dump4=50)
    assert len(lines) == 50
    assert sorted(lines) == first


def test_batch_equals_individual(eng):
    rng = random.Random(7)
    los, his = [], []
    for _ in range(24):
        lo = rng.randrange(1, 10**11)
        span = int(10 ** rng.uniform(2, 6))
        los.append(lo)
        his.append(lo + span)
    los.append(los[3]); his.append(his[3] + 1000)        # overlapping intervals
    los.append(5); his.append(5)                          # empty
    say("batch of 26 intervals")
    batch = eng.search_batch(los, his)
    for lo, hi, (res, lines) in zip(los, his, batch):
        one, l1 = eng.search(lo, hi)
        assert res == one, (lo, hi)
        assert sorted(lines) == sorted(l1), (lo, hi)


def test_partition_invariance(eng):
    lo, hi = 7 * 10**9, 7 * 10**9 + 3 * 10**7
    say("partition invariance")
    whole, lw = eng.search(lo, hi)
    cuts = [lo, lo + 123457, lo + 10**7 + 3, hi]
    parts = [eng.search(a, b) for a, b in zip(cuts, cuts[1:])]
    for k in whole["counters"]:
        assert whole["counters"][k] == sum(p[0]["counters"][k] for p in parts)
    assert sorted(lw) == sorted(sum((p[1] for p in parts), []))


def test_triple_queries_vs_oracle(eng):
    rng = random.Random(11)
    _, lines = ora.census(10**9, 10**9 + 3000, noext=True, dump4=2**62)
    _, lines2 = ora.census(10**5, 10**5 + 20000, noext=True, dump4=2**62)
    quads = [[int(x) for x in ln.split("[")[1].split("]")[0].split(",")] for ln in lines + lines2]
    rng.shuffle(quads)
    triples = []
    for v in quads[:40]:
        drop = rng.randrange(4)
        triples.append([x for i, x in enumerate(v) if i != drop])
    say(f"{len(triples)} triple queries")
    for d_max in (10**12, 10**8):
        got = eng.extend_triple_batch(triples, d_max)
        for abc, g in zip(triples, got):
            assert g == ora.extend_triple(abc, d_max), abc
    assert eng.extend_triple(triples[0], 10**12) == ora.extend_triple(triples[0], 10**12)
    with pytest.raises(ps4.PS4Error):
        eng.extend_triple([1, 2, 3], 100)


def test_capacity_failures_are_fatal():
    """Forced capacity overflows fail with PS4_E_CAPACITY, never truncate
    (the -DS4_MAX_* preflight of the reference, as runtime caps)."""
    for kw, lo, hi in (({"cap_divs": 64}, 1, 10**6), ({"cap_reps": 3}, 1, 10**6), ({"cap_ext": 1}, 1, 10**7)):
        say(f"capacity {kw}")
        e = ps4.Engine(S_hi_max=10**7, **kw)
        try:
            with pytest.raises(ps4.PS4Error) as ex:
                e.search(lo, hi)
            assert ex.value.rc == -2
        finally:
            e.close()


def test_verify_and_drill(golden):
    e = ps4.Engine(S_hi_max=10**7, verify=True, test6=True)
    try:
        e.set_alert(True)
        say("verify + test6 [1,1e7)")
        res, lines = e.search(1, 10**7)
        rec = next(r for r in golden["runs"] if (r["lo"], r["hi"]) == (1, 10**7))
        check(rec, res, lines)
        assert len(e.alerts) == 1
        assert e.alerts[0].startswith("TEST-6SET (NOT A SOLUTION, --test-sixset drill): [")
        assert e.alerts[0].rstrip("\n").endswith("*** VERIFY FAILED ***")
    finally:
        e.close()


def _run_cli(args, cwd):
    say("cli " + " ".join(str(a) for a in args))
    return ps4.run_cli(args, cwd=cwd, timeout=300)


def test_cli_drop_in(tmp_path, golden):
    """The ps4 CLI: s4 grammar, output file in append mode, stderr reports that the
    reference's own parsers accept (exhaustive/verify.py:103-135, compact.py:107-109,
    run_exhaustive.py:80-82)."""
    out = tmp_path / "o.out"
    r = _run_cli([1, 10**7, 4, 20, "--out", out, "--ckpt", tmp_path / "c.ckpt"], tmp_path)
    assert r.returncode == 0, r.stderr.decode()
    err = r.stderr.decode()
    assert re.search(r"^s4: S in \[1, 10000000\), 4 threads, chunk 1048576, ext=on, checkpointing$", err, re.M)
    rows = re.findall(r"^\s+thread\s+(\d+): S \[(\d+),(\d+)\)", err, re.M)
    assert [tuple(map(int, x)) for x in rows] == [(i, 1 + (10**7 - 1) * i // 4, 1 + (10**7 - 1) * (i + 1) // 4)
                                                 for i in range(4)]
    tot = re.findall(r"^TOTAL: (.*)$", err, re.M)
    assert len(tot) == 1
    kv = {k: int(v) for k, v in re.findall(r"([\w-]+)=(\d+)", tot[0])}
    rec = next(x for x in golden["runs"] if (x["lo"], x["hi"]) == (1, 10**7))
    for k, v in rec["counters"].items():
        assert kv[k] == v
    assert re.search(r"^CAPACITY: reps=24/512 divisors=252/16384 extensions=2/64$", err, re.M)
    hist = re.search(r"^5-sets by log2\(largest element\):(.*)$", err, re.M).group(1)
    assert dict(re.findall(r"(\d+):(\d+)", hist)) == {k: str(v) for k, v in rec["hist5"].items()}
    lines = out.read_text().splitlines()
    assert sha_sorted(lines) == rec["sorted_sha256"]
    ck = (tmp_path / "c.ckpt").read_text().splitlines()
    assert ck[0] == "s4ckpt v1" and ck[-1] == "DONE"
    # the audit chain's checks of a completed block (exhaustive/compact.py:29-45, 106-110):
    # checkpoint parameters, one 13-field row per thread covering its whole region, the
    # rows' counters summing to TOTAL, and the thread lines agreeing with the rows
    lo, hi, nt = 1, 10**7, 4
    assert tuple(map(int, ck[1].split())) == (lo, hi, nt, 1 << 20)
    assert len(ck) == nt + 3
    ckrows = [list(map(int, line.split())) for line in ck[2:-1]]
    for i, row in enumerate(ckrows):
        start, end = lo + (hi - lo) * i // nt, lo + (hi - lo) * (i + 1) // nt
        assert len(row) == 13 and row[:3] == [i, end, end] and row[3] == end - start and min(row) >= 0
    order = ("S_done", "pairs", "S_ge3", "triples", "4-sets", "5-sets", "6-sets", "multi-ext-4sets", "ext_cand",
             "ext_sqrt")
    for j, key in enumerate(order):
        assert sum(row[3 + j] for row in ckrows) == kv[key], key
    threads = re.findall(r"thread\s+(\d+): S \[(\d+),(\d+)\) 4-sets (\d+) 5-sets (\d+)", err)
    assert threads == [(str(i), str(lo + (hi - lo) * i // nt), str(row[2]), str(row[7]), str(row[8]))
                       for i, row in enumerate(ckrows)]
    # a finished checkpoint: --resume exits at once
    r2 = _run_cli([1, 10**7, 4, 20, "--out", out, "--ckpt", tmp_path / "c.ckpt", "--resume"], tmp_path)
    assert r2.returncode == 0 and b"nothing to do" in r2.stderr
    # an existing checkpoint without --resume is refused (exit 1)
    r3 = _run_cli([1, 10**7, 4, 20, "--ckpt", tmp_path / "c.ckpt"], tmp_path)
    assert r3.returncode == 1


def _total(err):
    return {k: int(v) for k, v in re.findall(r"([\w-]+)=(\d+)", re.search(r"^TOTAL: (.*)$", err, re.M).group(1))}


def test_cli_resume_equivalence(tmp_path, golden):
    """Interrupted-then-resumed equals one run.  The checkpoint is the state a kill
    after the first checkpoint leaves (A::sources/exhaustive/test_preflight.py:49-66):
    region 0 of 2 done up to a chunk boundary, region 1 untouched, results of the
    finished part already in the output file."""
    lo, hi, T, W = 1, 10**7, 2, 1 << 20
    half = lo + (hi - lo) // 2
    mid = lo + 3 * W
    part = _run_cli([lo, mid, 1, 20, "--out", tmp_path / "r.out"], tmp_path)
    assert part.returncode == 0
    done = _total(part.stderr.decode())
    ck = tmp_path / "r.ckpt"
    ck.write_text("s4ckpt v1\n%d %d %d %d\n0 %d %d %s\n1 %d %d %s\n" % (
        lo, hi, T, W, mid, half, " ".join(str(done[k]) for k in ps4.TOTAL_KEYS), half, hi, " ".join(["0"] * 10)))
    r = _run_cli([lo, hi, T, 20, "--out", tmp_path / "r.out", "--ckpt", ck, "--resume"], tmp_path)
    assert r.returncode == 0, r.stderr.decode()
    err = r.stderr.decode()
    assert "RESUMED from checkpoint" in err
    rec = next(x for x in golden["runs"] if (x["lo"], x["hi"]) == (1, 10**7))
    tot = _total(err)
    for k, v in rec["counters"].items():
        assert tot[k] == v, k
    assert sha_sorted((tmp_path / "r.out").read_text().splitlines()) == rec["sorted_sha256"]
    assert ck.read_text().splitlines()[-1] == "DONE"


def test_cli_fatal_exit_code(tmp_path):
    """Forced capacity failure through the CLI: FATAL line and exit code 2
    (exhaustive/s4.c:56-59; the reference forces it with -DS4_MAX_DIVS=1)."""
    env = dict(os.environ, PS4_CAP_DIVS="1")
    say("cli capacity failure")
    r = subprocess.run([ps4.CLI_PATH, "1", "1000000"], cwd=tmp_path, env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, timeout=300)
    assert r.returncode == 2
    assert b"FATAL: divisor capacity exceeded; this interval is NOT certified complete" in r.stderr


def test_cli_dump4(tmp_path):
    r = _run_cli([1, 10**6, 1, 20, "--noext", "--dump4", "5"], tmp_path)
    assert r.returncode == 0
    assert len([ln for ln in r.stdout.decode().splitlines() if ln.startswith("4:")]) == 5


def test_fixtures_1e8_1e9(eng, golden):
    for span in (10**8 - 1, 10**9 - 1):
        rec = next(r for r in golden["runs"] if (r["lo"], r["hi"]) == (1, 1 + span))
        say(f"fixture [1,{1 + span})")
        res, lines = eng.search(1, 1 + span)
        check(rec, res, lines)


def test_overflowing_batches_are_redone(eng, golden):
    """A batch whose 4-sets overflow the queue is redone smaller, never
    truncated, also when the next batch is already in flight: with the
    planner's 4-set estimate scaled to 5 % (a test hook) the first attempts
    overflow a 2^20-entry queue and are redone; [1,1e9) still equals s4's run
    over the same range, and uneven subranges in one batched call equal the
    main engine's."""
    rec = next(r for r in golden["runs"] if (r["lo"], r["hi"]) == (1, 10**9))
    from ps4 import workloads
    ranges = workloads.batched_ranges(64, seed=5, top=10**9, e0=4, e1=8)
    los, his = [a for a, _ in ranges], [b for _, b in ranges]
    ref_b = eng.search_batch(los, his)
    engines = []
    for pct in ("100", "5"):
        os.environ["PS4_QUAD_CAP"] = str(1 << 20)
        os.environ["PS4_TEST_DENSITY_PCT"] = pct
        try:
            engines.append(ps4.Engine(S_hi_max=10**9 + 1))
        finally:
            del os.environ["PS4_QUAD_CAP"], os.environ["PS4_TEST_DENSITY_PCT"]
    try:
        for e, pct in zip(engines, ("100", "5")):
            say(f"[1,1e9) with a 2^20-entry queue, estimate at {pct} %")
            e.kernel_time(reset=True)
            res, lines = e.search(1, 10**9)
            check(rec, res, lines)
            say(f"uneven subranges, estimate at {pct} %")
            for r, (c, ls), (c0, l0) in zip(ranges, e.search_batch(los, his), ref_b):
                assert c == c0 and sorted(ls) == sorted(l0), (r, pct)
        assert engines[0].redo_count() == 0 and engines[1].redo_count() > 0
        assert engines[0].kernel_time()[2] > 1
    finally:
        for e in engines:
            e.close()


def test_high_s_fixtures(eng, golden):
    for rec in golden["runs"]:
        if rec["hi"] - rec["lo"] > 10**5 and rec["lo"] >= 9 * 10**11:
            say(f"fixture [{rec['lo']},{rec['hi']})")
            res, lines = eng.search(rec["lo"], rec["hi"])
            check(rec, res, lines, (rec["lo"], rec["hi"]))


def test_census_1e10(eng, golden):
    """The headline range [1,1e10): every counter and all 54477 lines."""
    rec = next(r for r in golden["runs"] if (r["lo"], r["hi"]) == (1, 10**10))
    say("[1,1e10)")
    res, lines = eng.search(1, 10**10)
    check(rec, res, lines)
    # its M: lines' primitive near misses, against the reference's catalogs
    from ps4 import nearmiss
    near, sols = nearmiss.classify(lines)
    assert not sols and len(near) > 18
    assert _near_misses_in_reference_catalogs(near) > 0


def test_census_first_thread_range_of_1e11_log(eng, golden):
    """[1, 6.25e9), the first thread's range of exhaustive/logs/s4d_1e11.err:
    every counter and all 38019 lines against exhaustive/s4.c."""
    rec = next(r for r in golden["runs"] if (r["lo"], r["hi"]) == (1, 6250000000))
    say("[1,6.25e9)")
    res, lines = eng.search(1, 6250000000)
    check(rec, res, lines)


def test_census_1e11_reference_log(eng, golden):
    """[1,1e11) against the reference's own log (exhaustive/logs/s4d_1e11.err:68-69).

    The log was written by an earlier build of the program than
    exhaustive/s4.c: its TOTAL line carries "nonprimitive skipped=", which s4.c
    does not print, and on the log's first thread range [1, 6.25e9) s4.c finds
    759664239 4-sets and 37818 5-sets where the log has 759675299 and 37820
    (the golden run of that range).  The counters up to the representation
    triples do not depend on that difference and must agree; the 4-set and
    extension counters are checked against s4.c itself on [1, 6.25e9) and
    [1, 1e10)."""
    r = next(x for x in golden["reference_logs"] if x["source"].endswith("s4d_1e11.err"))
    say("[1,1e11)")
    res, lines = eng.search(r["lo"], r["hi"])
    for k in ("S_done", "pairs", "S_ge3", "triples", "6-sets"):
        assert res["counters"][k] == r["counters"][k], k
    assert sum(ln.startswith("5:") for ln in lines) == res["counters"]["5-sets"]
    assert sum(ln.startswith("M:") for ln in lines) == res["counters"]["multi-ext-4sets"]
    assert sum(res["hist5"].values()) == res["counters"]["5-sets"]


def test_search_batch_uneven_subranges(eng):
    """Seeded uneven subranges (the batched workload's generator, smaller): one
    batched call equals one call per subrange, counters and lines."""
    from ps4 import workloads
    ranges = workloads.batched_ranges(48, seed=42, top=10**10, e0=3, e1=6)
    say(f"batched {len(ranges)} subranges")
    got = eng.search_batch([a for a, _ in ranges], [b for _, b in ranges])
    for (a, b), (c, lines) in zip(ranges, got):
        alone = eng.search(a, b)
        assert alone[0] == c, (a, b)
        assert sorted(alone[1]) == sorted(lines), (a, b)


def test_tuning_never_changes_results(eng):
    """Launch shape and batch plan are timing only: extreme tunings (one
    workgroup per CU, every compiled k_rq block size and tile, every k_ext
    claim size, 8x finer subrange split, tiny batches) give the default
    results on a census and on uneven subranges; values above the maxima and
    shapes that are not compiled are rejected."""
    from ps4 import workloads
    base, mx = eng.get_tuning()
    n_cu = mx["n_cu"]
    # (the maximum grid is that of the compiled shape with the most resident workgroups)
    assert n_cu > 0 and 0 < base["rq_grid"] <= mx["rq_grid"] and base["batch_pct"] == 80
    assert (base["rq_block"], base["rq_tile"], base["ext_chunk"]) == (256, 4, 4)
    ranges = workloads.batched_ranges(24, seed=3, top=10**11, e0=3, e1=7)
    los, his = [a for a, _ in ranges], [b for _, b in ranges]
    say("tuning: default results")
    ref_c = eng.search(2 * 10**9, 2 * 10**9 + 2 * 10**7)
    ref_b = eng.search_batch(los, his)
    tunings = [
        {"rq_grid": n_cu, "ext_grid": n_cu, "rq_block": 128, "rq_tile": 2, "ext_chunk": 1, "sw_shift": 3,
         "batch_pct": 5},
        {"rq_grid": mx["rq_grid"] - n_cu, "ext_grid": mx["ext_grid"] - n_cu, "rq_block": 512, "rq_tile": 8,
         "ext_chunk": 16, "sw_shift": 1, "batch_pct": 95},
        {"rq_block": 256, "rq_tile": 8, "ext_chunk": 1},
    ]
    try:
        for t in tunings:
            say(f"tuning: {t}")
            eng.set_tuning(t)
            cur, _ = eng.get_tuning()
            assert all(cur[k] == v for k, v in t.items())
            assert eng.search(2 * 10**9, 2 * 10**9 + 2 * 10**7) == ref_c
            got = eng.search_batch(los, his)
            for r, (c, lines), (c0, l0) in zip(ranges, got, ref_b):
                assert c == c0, (r, t)
                assert sorted(lines) == sorted(l0), (r, t)
        for over in ({"rq_grid": mx["rq_grid"] + 1}, {"ext_grid": mx["ext_grid"] + 1}, {"sw_shift": 9},
                     {"batch_pct": 96}, {"rq_block": 64}, {"rq_tile": 3}, {"ext_chunk": 2}):
            with pytest.raises(ps4.PS4Error):
                eng.set_tuning(over)
    finally:
        eng.set_tuning({})
    assert eng.get_tuning()[0] == base


def test_autotune_runs_and_keeps_results(eng, tmp_path):
    """The autotuner on a small sample: every trial's results must equal the
    default's (it raises otherwise); its file reads back and applies."""
    from ps4 import autotune
    say("autotune on 16 subranges")
    try:
        best, rep = autotune.tune(eng, n=16, seed=5, top=10**10, e0=3, e1=6, reps=1, log=say)
        assert rep["trials"] and all(t["seconds"] > 0 for t in rep["trials"])
        f = tmp_path / "tune.txt"
        f.write_text(ps4.format_tuning(best, "test"))
        eng.set_tuning(str(f))
        cur, _ = eng.get_tuning()
        assert {k: cur[k] for k in best} == best
    finally:
        eng.set_tuning({})


def test_cli_tune_file(tmp_path, golden):
    """ps4 --tune FILE: same output as without; a bad file is a usage error."""
    good = tmp_path / "t.txt"
    good.write_text("sw_shift = 2\nbatch_pct = 30\n")
    r = _run_cli([1, 10**7, "--tune", good, "--out", tmp_path / "o.out"], tmp_path)
    assert r.returncode == 0, r.stderr.decode()
    rec = next(x for x in golden["runs"] if (x["lo"], x["hi"]) == (1, 10**7))
    assert sha_sorted((tmp_path / "o.out").read_text().splitlines()) == rec["sorted_sha256"]
    bad = tmp_path / "b.txt"
    bad.write_text("threads = 4\n")
    r = _run_cli([1, 10**6, "--tune", bad], tmp_path)
    assert r.returncode == 1


def test_graph_launches_equal_eager(eng):
    """PS4_GRAPH (each batch's copy, reset, kernels and read-backs as one captured
    hipGraph, updated in place between batches): every replay equals the eager
    engine's result, with eager work and synchronisations between the replays;
    a census of several batches (small batch fill) and uneven subranges."""
    from ps4 import workloads
    ranges = workloads.batched_ranges(24, seed=11, top=10**11, e0=3, e1=7)
    los, his = [a for a, _ in ranges], [b for _, b in ranges]
    say("graph: eager references")
    ref_c = eng.search(3 * 10**9, 3 * 10**9 + 3 * 10**7)
    ref_b = eng.search_batch(los, his)
    os.environ["PS4_QUAD_CAP"] = str(1 << 22)   # a small 4-set queue: several batches per call
    try:
        g = ps4.Engine(S_hi_max=10**11 + 1, graph=True)
    finally:
        del os.environ["PS4_QUAD_CAP"]
    try:
        g.set_tuning({"batch_pct": 30})         # the graph is updated between batches
        for rep in range(3):
            say(f"graph: replay {rep}")
            assert g.search(3 * 10**9, 3 * 10**9 + 3 * 10**7) == ref_c
            got = g.search_batch(los, his)
            for r, (c, lines), (c0, l0) in zip(ranges, got, ref_b):
                assert c == c0, (r, rep)
                assert sorted(lines) == sorted(l0), (r, rep)
            assert eng.search(10**6, 2 * 10**6)[0]["counters"]["S_done"] == 10**6   # eager work between
        assert g.kernel_time()[2] >= 12   # several graph launches per call
        g.set_tuning({})
        res, lines = g.search(1, 10**8)
        one, l1 = eng.search(1, 10**8)
        assert res == one and sorted(lines) == sorted(l1)
    finally:
        g.close()


def test_near_misses_of_the_census_mlines(eng, golden):
    """Every M: line of the [1,10^9) census (59, as in the reference's run) splits
    into near misses with 14 of 15 square sums (checked inside split_mline) and
    no 6-set solution (the reference finds none below 10^13); scaled copies
    share their primitive forms (18 distinct ones here), each of them one of
    the reference's primitive near misses with the same pair and Euler count."""
    from ps4 import nearmiss
    say("near misses of [1,1e9) M-lines")
    res, lines = eng.search(1, 10**9)
    mlines = [x for x in lines if x.startswith("M:")]
    assert len(mlines) == res["counters"]["multi-ext-4sets"] == 59
    near, sols = nearmiss.classify(mlines)
    # M-lines that are square multiples of one another share their primitive near misses
    pairs = sum(math.comb(len(nearmiss.parse_mline(x)[1]), 2) for x in mlines)
    assert not sols and 0 < len(near) <= pairs
    assert all(nearmiss.split_mline(x)[1] for x in mlines)      # each line gives a near miss
    for prim, (pair, eu, S) in near.items():
        assert set(pair) <= set(prim) and S < 10**9
    _near_misses_in_reference_catalogs(near)


def _near_misses_in_reference_catalogs(near):
    """Each primitive near miss is one of the reference's primitive near misses
    below S = 1e12 (exhaustive/analysis_1e12.txt) with the same non-square
    pair, and where the reference's catalogs record its Euler count
    (exhaustive/evidence/*-compact.tar.gz::*/near-misses.json) it is the same."""
    import json
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "near_misses.json")) as f:
        d = json.load(f)
    below = {tuple(e["values"]): tuple(e["nonsquare_pair"]) for e in d["below_1e12"]}
    euler = {tuple(e["values"]): e["euler_squares"] for e in d["catalog"]}
    n_eu = 0
    for prim, (pair, eu, _S) in near.items():
        assert below.get(prim) == tuple(pair), prim
        if prim in euler:
            assert eu == euler[prim], prim
            n_eu += 1
    return n_eu


def test_serving_results_same_alone_and_in_the_mix(eng):
    """Requests through the continuous-batching loop (arriving while others run)
    give exactly the results of the same requests run alone."""
    from ps4 import serve, workloads
    stream = workloads.serving_stream(80, rate=2000.0, seed=7, s_top=10**10, len_e0=3, len_e1=6, c_max=10**7,
                                      d_max=10**10)
    say(f"serving {len(stream)} requests")
    srv = serve.Server(eng, step_quads=1 << 14)   # small steps: requests are cut into several parts
    t0 = time.perf_counter()
    reqs = []
    for t_arr, kind, a in stream:
        dt = t0 + t_arr - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        reqs.append(srv.submit(kind, a))
    for r in reqs:
        assert r.wait(120)
    srv.close()
    assert any(n > 1 for n, _ in srv.steps)     # some steps finished several requests
    for r in reqs:
        assert r.error is None
        if r.kind == "range":
            alone = eng.search(*r.args)
            assert alone[0] == r.result[0] and sorted(alone[1]) == sorted(r.result[1]), r.args
        else:
            assert eng.extend_triple(*r.args) == r.result, r.args


def _closed_form_square_triples(n_max):
    """Square triples {a<b<c}, c <= n_max, counted per root pair (x, y) in closed form."""
    n = 0
    x = 1
    while x * x <= 2 * n_max:
        y = x + 1
        while 2 * y + 3 <= x * x and y * y + (y + 1) ** 2 <= 2 * n_max + x * x:
            zmax = min(math.isqrt(x * x + y * y - 2), math.isqrt(2 * n_max + x * x - y * y))
            zl = y + 1 + ((y + 1 + x + y) & 1)
            if zmax >= zl:
                n += (zmax - zl) // 2 + 1
            y += 1
        x += 1
    return n


def test_elements_small_vs_oracle():
    """Element-range search for small N (edge cases included) against the
    brute-force oracle: the 4-set lines and the square-triple count."""
    e = ps4.Engine(S_hi_max=4 * 10**5)
    try:
        for n_max in (1, 3, 4, 7, 100, 1000, 25000, 10**5):
            say(f"elements n_max={n_max}")
            lines, ntri = e.search_elements(n_max)
            assert lines == ora.elements(n_max), n_max
            assert ntri == ora.square_triples(n_max), n_max
    finally:
        e.close()


def test_elements_reference_fixtures(golden):
    """N = 25000, 1e6, 1e7: the 4-sets with largest element <= N from the
    reference's --dump4 output (count and SHA-256 of the sorted lines); the
    square-triple count against the oracle's enumeration of every root triple
    (tests/golden/make_golden.py --square-triples), and for N <= 1e6 also
    against a per-root-pair closed form."""
    e = ps4.Engine(S_hi_max=4 * 10**7)
    try:
        for rec in golden["elements"]:
            say(f"elements n_max={rec['n_max']}")
            lines, ntri = e.search_elements(rec["n_max"])
            assert len(lines) == rec["n_sets"]
            assert hashlib.sha256("".join(x + "\n" for x in lines).encode()).hexdigest() == rec["sha256"]
            # enumerated root triples (the oracle's direct count, recorded in the fixture)
            assert ntri == rec["square_triples"]
            if rec["n_max"] <= 10**6:
                assert ntri == _closed_form_square_triples(rec["n_max"])
    finally:
        e.close()


def test_elements_limits():
    """n_max beyond the tables is an argument error, not a truncated answer."""
    e = ps4.Engine(S_hi_max=10**6)
    try:
        with pytest.raises(RuntimeError):
            e.search_elements(10**6)
    finally:
        e.close()


def test_cli_elements(tmp_path):
    say("cli --elements 25000")
    exe = os.path.join(ps4.PKG_DIR, "lib", "ps4")
    r = subprocess.run([exe, "--elements", "25000"], stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.decode().splitlines()
    assert lines == ora.elements(25000)
    assert f"ELEMENTS: n_max=25000 sets={len(lines)} square_triples={ora.square_triples(25000)}" in r.stderr.decode()

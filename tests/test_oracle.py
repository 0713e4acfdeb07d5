"""Pins of the CPU restatement (oracle/cpu_ref.c) -- CPU only.

The oracle is checked against what the reference itself fixes:
  * the counter vectors the reference ships in exhaustive/logs/*.err,
  * the known answers of exhaustive/README.md:157-159, exhaustive/test_verify.py:13,31-32
    and bruteforce-scratchpad/dotnet/1mln.txt,
  * fixtures written by the reference s4.c (tests/golden/make_golden.py),
and against mathematics where those leave gaps: brute-force enumeration of
4-sets and extensions on tiny inputs, the closed form r^2 <= n < (r+1)^2,
divisor counts by definition.
"""
import hashlib
import math
import random

import pytest

import ora
from conftest import sha_sorted

FAST_SPAN = 2 * 10**6


def _check_run(rec, res, lines):
    assert res["counters"] == rec["counters"]
    assert res["hist5"] == rec["hist5"]
    if rec.get("capacity"):
        assert res["capacity"] == rec["capacity"]
    assert len(lines) == rec["n_lines"]
    assert sha_sorted(lines) == rec["sorted_sha256"]
    if "lines" in rec:
        assert sorted(lines) == sorted(rec["lines"])


def test_fixture_runs_small(golden):
    """Every reference run up to FAST_SPAN wide: counters, histogram, capacity, lines."""
    n = 0
    for rec in golden["runs"]:
        if rec["hi"] - rec["lo"] > FAST_SPAN or rec["lo"] > 2 * 10**12:
            continue
        res, lines = ora.census(rec["lo"], rec["hi"])
        _check_run(rec, res, lines)
        n += 1
    assert n >= 25


def test_fixture_1e7_and_frontier(golden):
    """[1,1e7) (README known answer) and the highest fixture window near 1.18e13."""
    want = {(1, 10**7), (11762800000000, 11762800300000)}
    seen = 0
    for rec in golden["runs"]:
        if (rec["lo"], rec["hi"]) in want:
            res, lines = ora.census(rec["lo"], rec["hi"])
            _check_run(rec, res, lines)
            seen += 1
    assert seen == 2


def test_readme_known_answer():
    """exhaustive/README.md:157-159: [1,1e7) has 123 five-sets, 2 M-lines, 0 sextuples."""
    res, lines = ora.census(1, 10**7)
    c = res["counters"]
    assert (c["5-sets"], c["multi-ext-4sets"], c["6-sets"]) == (123, 2, 0)
    assert sum(ln.startswith("5:") for ln in lines) == 123
    assert sum(ln.startswith("M:") for ln in lines) == 2


def test_reference_log_singletons(golden):
    """The nine boundary repairs S = k*1e11 (exhaustive/logs/boundary_k.err:3)."""
    recs = [r for r in golden["reference_logs"] if r["source"].startswith("exhaustive/logs/boundary_")]
    assert len(recs) == 9
    quads = 0
    for r in recs:
        res, lines = ora.census(r["lo"], r["hi"])
        assert res["counters"] == r["counters"], r["source"]
        quads += res["counters"]["4-sets"]
    assert quads == 15                                  # exhaustive/README.md:94


def _mline(line):
    head, ext = line.split(" ext=[")
    return head, sorted(int(x) for x in ext.split("]")[0].split(","))


def test_known_lines(golden):
    """The auditor's known lines.  exhaustive/test_verify.py:31-32 lists the three
    extensions in another order than s4.c prints them (its checks treat them as a
    set), so that line is compared as members + extension set; the exact s4.c
    bytes for that S are in the reference-generated fixtures."""
    kl = golden["known_lines"]
    _, lines = ora.census(362141, 362142)
    assert kl["exhaustive/test_verify.py:13"] in lines
    _, lines = ora.census(50578642000, 50578642001)
    ms = [_mline(ln) for ln in lines if ln.startswith("M:")]
    assert _mline(kl["exhaustive/test_verify.py:31-32"]) in ms


def test_known_five_sets_below_1e6(golden):
    """The only two 5-sets with largest element < 1e6 (1mln.txt) are in the census."""
    _, lines = ora.census(1, 4 * 10**6)
    fives = set()
    for ln in lines:
        if ln.startswith("5:"):
            vals = tuple(sorted(int(x) for x in ln.split("[")[1].split("]")[0].split(",")))
            if max(vals) < 10**6:
                fives.add(vals)
    assert fives == {tuple(v) for v in golden["known_5sets_below_1e6"]}


def test_dump4_fixtures(golden):
    for rec in golden["dump4"]:
        if rec["hi"] - rec["lo"] > FAST_SPAN:
            continue
        res, lines = ora.census(rec["lo"], rec["hi"], noext=True, dump4=2**62)
        assert res["counters"] == rec["counters"]
        assert sha_sorted(lines) == rec["sorted_sha256"]


def _brute_quads(S_max):
    """All {a>b>c>d>0} with six square pair sums and a+b+c+d < S_max, by definition."""
    sq = set(i * i for i in range(1, int(math.isqrt(S_max)) + 2))
    out = set()
    for a in range(1, S_max):
        for b in range(1, min(a, S_max - a)):
            if a + b not in sq:
                continue
            for c in range(1, min(b, S_max - a - b)):
                if a + c not in sq or b + c not in sq:
                    continue
                for d in range(1, min(c, S_max - a - b - c)):
                    if a + d in sq and b + d in sq and c + d in sq:
                        out.add((a, b, c, d))
    return out


def test_four_sets_brute_force():
    """Every 4-set with S < 6000 exactly once, against direct enumeration."""
    S_max = 6000
    _, lines = ora.census(1, S_max, noext=True, dump4=2**62)
    got = [tuple(int(x) for x in ln.split("[")[1].split("]")[0].split(",")) for ln in lines]
    assert len(got) == len(set(got))
    assert set(got) == _brute_quads(S_max)


def _brute_ext(v):
    """Extensions e of the 4-set v by direct search over e (small members only)."""
    out = []
    top = 4 * max(v) ** 2
    e = 1
    while e < top:
        if e not in v and all(math.isqrt(e + x) ** 2 == e + x for x in v):
            out.append(e)
        e += 1
    return out


def test_extension_brute_force():
    """5-sets from the census equal a direct search for e on a few small 4-sets."""
    # 4-sets of S = 4205 and S = 5525 with their complete extension lists
    for S in (4205, 5525, 7585):
        _, q4 = ora.census(S, S + 1, noext=True, dump4=2**62)
        _, full = ora.census(S, S + 1)
        fives = {tuple(sorted(int(x) for x in ln.split("[")[1].split("]")[0].split(","))) for ln in full
                 if ln.startswith("5:")}
        for ln in q4:
            v = [int(x) for x in ln.split("[")[1].split("]")[0].split(",")]
            if max(v) > 3000:
                continue
            ext = [e for e in _brute_ext(v) if e <= 10**6]
            for e in ext:
                if e > max(v):
                    assert tuple(sorted(v + [e])) in fives


def test_isqrt_closed_form():
    rng = random.Random(1)
    for n in list(range(0, 1 << 16)) + [rng.randrange(1 << 64) for _ in range(20000)]:
        r = ora.isqrt64(n)
        assert r * r <= n < (r + 1) * (r + 1)
    for k in (2**32 - 1, 2**32, 3037000499, 2**42 + 7, 2**43 - 1):
        for n in (k * k - 1, k * k, k * k + 1):
            r = ora.isqrt128(n)
            assert r * r <= n < (r + 1) ** 2
            assert ora.is_square128(n) == (n == k * k)
    for _ in range(3000):
        n = rng.randrange(1 << 100)
        r = ora.isqrt128(n)
        assert r * r <= n < (r + 1) ** 2
        assert ora.is_square128(r * r) and not ora.is_square128(r * r + 1)


def test_divisor_counts_and_order():
    for n in range(1, 3000):
        assert ora.num_divisors(n) == sum(1 for d in range(1, n + 1) if n % d == 0)
    rng = random.Random(2)
    for _ in range(300):
        m1, m2 = rng.randrange(1, 5000), rng.randrange(1, 5000)
        D = m1 * m2
        ds = ora.divisors_of_product(m1, m2)
        assert sorted(ds) == [d for d in range(1, D + 1) if D % d == 0] if D < 200000 else True
        assert len(ds) == ora.num_divisors(D)
        for t, d in enumerate(ds):          # mixed-radix complement
            assert d * ds[len(ds) - 1 - t] == D


def _brute_triple(abc, d_max):
    a, b, c = abc
    return [d for d in range(1, d_max + 1) if d not in abc and all(math.isqrt(d + x) ** 2 == d + x for x in abc)]


def test_triple_query_brute_force():
    """Triple -> d on small square triples against a direct scan over d."""
    _, lines = ora.census(1, 12000, noext=True, dump4=2**62)
    seen = 0
    for ln in lines:
        v = [int(x) for x in ln.split("[")[1].split("]")[0].split(",")]
        for drop in range(4):
            abc = [x for i, x in enumerate(v) if i != drop]
            d_max = 20000
            got = ora.extend_triple(abc, d_max)
            assert got == _brute_triple(abc, d_max)
            assert v[drop] in got
            seen += 1
    assert seen > 0


def test_triple_query_matches_dump4():
    """Every 4-set of a --dump4 window is found back from each of its triples."""
    _, lines = ora.census(10**6, 10**6 + 20000, noext=True, dump4=2**62)
    for ln in lines[:200]:
        v = [int(x) for x in ln.split("[")[1].split("]")[0].split(",")]
        for drop in range(4):
            abc = [x for i, x in enumerate(v) if i != drop]
            assert v[drop] in ora.extend_triple(abc, 10**12)


def test_partition_invariance():
    """Counters and lines are additive over a split of the range (exhaustive/s4.c:451-456)."""
    lo, mid, hi = 3 * 10**6, 3 * 10**6 + 123457, 3 * 10**6 + 300000
    a, la = ora.census(lo, mid)
    b, lb = ora.census(mid, hi)
    w, lw = ora.census(lo, hi)
    for k in w["counters"]:
        assert w["counters"][k] == a["counters"][k] + b["counters"][k]
    assert sorted(la + lb) == lw


def test_planted_mistake_is_caught(golden):
    """A wrong counter vector fails the fixture comparison (the check has teeth)."""
    rec = next(r for r in golden["runs"] if (r["lo"], r["hi"]) == (1, 10**5))
    res, lines = ora.census(1, 10**5)
    bad = dict(res)
    bad["counters"] = dict(res["counters"], ext_sqrt=res["counters"]["ext_sqrt"] + 1)
    with pytest.raises(AssertionError):
        _check_run(rec, bad, lines)


# ---- element-range mode (the search by largest element) ----

def _closed_form_square_triples(n_max):
    """Square triples {a<b<c}, c <= n_max, counted per root pair (x, y) with the
    admissible z (y < z, x + y + z even, a >= 1, c <= n_max) in closed form."""
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


def test_square_triples_brute_force():
    """Every {a<b<c} with c <= 300 tried directly against the oracle's count."""
    sq = {i * i for i in range(60)}
    for n_max in (1, 10, 60, 150, 300):
        want = sum(1 for c in range(1, n_max + 1) for b in range(1, c) if b + c in sq
                   for a in range(1, b) if a + b in sq and a + c in sq)
        assert ora.square_triples(n_max) == want, n_max


def test_square_triples_closed_form():
    for n_max in (1000, 25000, 10**5):
        assert ora.square_triples(n_max) == _closed_form_square_triples(n_max), n_max


def test_elements_match_reference_dump(golden):
    """The oracle's element-range search equals the 4-sets with largest element
    <= N that the reference prints with --dump4 (tests/golden/make_golden.py
    --elements), and the [1,1e5) dump filtered by hand."""
    for rec in golden["elements"]:
        if rec["n_max"] > 10**5:
            continue   # larger ones are checked on the GPU
        lines = ora.elements(rec["n_max"])
        assert len(lines) == rec["n_sets"]
        assert hashlib.sha256("".join(x + "\n" for x in lines).encode()).hexdigest() == rec["sha256"]
    dump = next(r for r in golden["dump4"] if (r["lo"], r["hi"]) == (1, 10**5))
    n_max = 25000   # every 4-set with d <= 25000 has S <= 99994 < 1e5
    sets = sorted(sorted(int(v) for v in ln.split("[")[1].split("]")[0].split(",")) for ln in dump["lines"])
    want = ["[%d,%d,%d,%d]" % tuple(v) for v in sets if v[3] <= n_max]
    assert ora.elements(n_max) == want


def test_elements_are_square_sets():
    """Every set the oracle returns has six square pair sums and d <= N, and
    contains square triples counted by the triple count."""
    for ln in ora.elements(10**5):
        v = [int(x) for x in ln[1:-1].split(",")]
        assert v == sorted(v) and len(set(v)) == 4 and v[3] <= 10**5
        for i in range(4):
            for j in range(i + 1, 4):
                r = math.isqrt(v[i] + v[j])
                assert r * r == v[i] + v[j]

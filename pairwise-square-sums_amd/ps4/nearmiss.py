"""Near-miss post-processing of M: lines (host side, no device work).

An M: line of the census is a 4-set {a,b,c,d} with two or more extensions e,
each with e+a, e+b, e+c, e+d square.  Every pair of extensions (e, f) gives a
6-set whose fifteen pair sums are square except possibly e+f:
  - e+f square: a solution (the census also reports it as a 6-set);
  - otherwise a near miss with 14 of 15 square pair sums.  It is reduced to its
    primitive form by dividing every member by the largest square dividing
    their gcd, since scaling by a square keeps every square sum square.

The split and the scaling follow the reference's continuation check
(exhaustive/evidence/verify_continuation.original.py:34-60).  The Euler count
restates its euler_count (:62-74).  For each member t outside the non-square
pair {e, f}, the six members split into x = (e, f, t) and y = the other three.
The roots of the nine x-y sums form a 3x3 array r[i][j] = isqrt(x[(i+j)%3] +
y[(i+2j)%3]).  The split counts if some signs on rows 1 and 2 make the three
rows pairwise orthogonal.

Pinned by tests/test_workloads.py:
  - the reference's known M-line (exhaustive/test_verify.py:31-32), whose
    primitives appear with the same non-square pairs among the 430 primitive
    near misses of exhaustive/analysis_1e12.txt;
  - the Euler count of all 865 distinct entries of the reference's near-miss
    catalogs (exhaustive/evidence/*-compact.tar.gz::*/near-misses.json,
    field euler_squares), and their non-square pairs and square totals;
  - brute force of the square part;
  - the 14-of-15 count of every primitive.
"""
import itertools
import math
import re
from functools import reduce

_BRACKETS = re.compile(r"\[([0-9,\s]*)\]")


def is_square(n):
    r = math.isqrt(n)
    return r * r == n


def square_part(n):
    """Largest square dividing n > 0.  Primes up to n^(1/3) are divided out; the
    cofactor then has at most two prime factors, and adds a square only if it is one."""
    out, m, p = 1, n, 2
    while p * p * p <= m:
        while m % (p * p) == 0:
            m //= p * p
            out *= p * p
        if m % p == 0:
            m //= p
        p += 1 if p == 2 else 2
    return out * m if m > 1 and is_square(m) else out


def parse_mline(line):
    """(four members, extensions, S) of an 'M: [a,b,c,d] ext=[...] S=...' line."""
    groups = _BRACKETS.findall(line)
    if not line.startswith("M:") or len(groups) != 2:
        raise ValueError(f"not an M: line: {line!r}")
    four, ext = ([int(v) for v in g.split(",")] for g in groups)
    m = re.search(r"S=(\d+)", line)
    if len(four) != 4 or len(ext) < 2 or not m:
        raise ValueError(f"not an M: line: {line!r}")
    return four, ext, int(m.group(1))


def split_mline(line):
    """(S, {primitive 6-tuple: primitive non-square pair}, set of 6-set solutions)."""
    four, ext, S = parse_mline(line)
    if sum(four) != S or not all(is_square(u + v) for u, v in itertools.combinations(four, 2)) or \
            not all(is_square(e + u) for e in ext for u in four):
        raise ValueError(f"M: line does not check: {line!r}")
    near, solutions = {}, set()
    for e, f in itertools.combinations(ext, 2):
        six = four + [e, f]
        if is_square(e + f):
            solutions.add(tuple(sorted(six)))
            continue
        k = square_part(reduce(math.gcd, six))
        prim = tuple(sorted(v // k for v in six))
        n_sq = sum(is_square(u + v) for u, v in itertools.combinations(prim, 2))
        if n_sq != 14:
            raise AssertionError(f"near miss with {n_sq} square sums: {prim}")
        near[prim] = tuple(sorted((e // k, f // k)))
    return S, near, solutions


def euler_count(six, bad):
    """Splits of a primitive near miss (sorted 6-tuple six, non-square pair bad)
    whose 3x3 array of cross roots has pairwise orthogonal rows up to row signs."""
    count = 0
    for t in six:
        if t in bad:
            continue
        x = tuple(bad) + (t,)
        y = tuple(v for v in six if v not in x)
        r = [[math.isqrt(x[(i + j) % 3] + y[(i + 2 * j) % 3]) for j in range(3)] for i in range(3)]
        for s in itertools.product((1, -1), repeat=6):
            r1 = [r[1][j] * s[j] for j in range(3)]
            r2 = [r[2][j] * s[3 + j] for j in range(3)]
            if sum(a * b for a, b in zip(r[0], r1)) == 0 and sum(a * b for a, b in zip(r[0], r2)) == 0 and \
                    sum(a * b for a, b in zip(r1, r2)) == 0:
                count += 1
                break
    return count


def classify(lines):
    """Every near miss of the M: lines among lines: {primitive: (pair, euler count, S)},
    and the set of 6-set solutions."""
    near_all, sols = {}, set()
    for line in lines:
        if not line.startswith("M:"):
            continue
        S, near, solutions = split_mline(line)
        sols |= solutions
        for prim, pair in near.items():
            near_all.setdefault(prim, (pair, euler_count(prim, pair), S))
    return near_all, sols

"""Synthetic request mixes for the census engine.

The reference runs one interval per process (exhaustive/s4.c:416-426, driven
by run_exhaustive.py block by block).  The engine also takes many intervals at
once and triple queries; these generators make the fixed, seeded inputs that
measure those forms:

  batched_ranges  n subranges [lo, lo+len) of [1, top]: len = floor(10^(e0 +
                  (e1-e0)*U)), lo = 1 + floor(U' * (top - len)), U and U' from
                  splitmix64(seed) divided by 2^64 (two draws per range).
  band_ranges     a band cut into equal consecutive subranges.
  serving_stream  an open-loop Poisson stream of range queries and
                  triple-extension queries.
"""
import math
import random

MASK64 = (1 << 64) - 1


def splitmix64(seed):
    """The splitmix64 generator (Steele, Lea, Flood 2014): an endless iterator of u64."""
    state = seed & MASK64
    while True:
        state = (state + 0x9E3779B97F4A7C15) & MASK64
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        yield z ^ (z >> 31)


def batched_ranges(n=4096, seed=42, top=10**11, e0=4, e1=9):
    """n subranges of [1, top] with log-uniform lengths in [10^e0, 10^e1]."""
    g = splitmix64(seed)
    out = []
    for _ in range(n):
        u = next(g) / 2.0**64
        u2 = next(g) / 2.0**64
        length = int(10 ** (e0 + (e1 - e0) * u))
        lo = 1 + int(u2 * (top - length))
        out.append((lo, lo + length))
    return out


def band_ranges(lo, hi, width):
    """[lo, hi) cut into consecutive subranges of the given width (the last may be shorter)."""
    return [(a, min(hi, a + width)) for a in range(lo, hi, width)]


def random_square_triple(rng, c_max):
    """A triple a < b < c <= c_max of positive integers with a+b, a+c, b+c square,
    from roots x < y < z with x^2 + y^2 + z^2 even: a = (x^2+y^2-z^2)/2,
    b = (x^2-y^2+z^2)/2, c = (-x^2+y^2+z^2)/2 (then a+b = x^2, a+c = y^2,
    b+c = z^2)."""
    z_max = math.isqrt(2 * c_max)
    while True:
        z = rng.randrange(3, z_max + 1)
        y = rng.randrange(2, z)
        x = rng.randrange(1, y)
        s = x * x + y * y + z * z
        if s & 1 or x * x + y * y <= z * z:
            continue
        a, b, c = (x * x + y * y - z * z) // 2, (x * x - y * y + z * z) // 2, (-x * x + y * y + z * z) // 2
        if 0 < a < b < c <= c_max:
            return a, b, c


def serving_stream(n, rate=200.0, seed=7, frac_range=0.7, s_top=10**12, len_e0=3, len_e1=8, c_max=10**9,
                   d_max=10**12):
    """n requests of an open-loop Poisson stream: (arrival_s, kind, args) with
    kind "range" (args = (lo, hi), lo uniform in [1, s_top], length log-uniform
    in [10^len_e0, 10^len_e1]) or "triple" (args = ((a, b, c), d_max))."""
    rng = random.Random(seed)
    t = 0.0
    out = []
    for _ in range(n):
        t += rng.expovariate(rate)
        if rng.random() < frac_range:
            length = int(10 ** rng.uniform(len_e0, len_e1))
            lo = rng.randrange(1, s_top)
            out.append((t, "range", (lo, lo + length)))
        else:
            out.append((t, "triple", (random_square_triple(rng, c_max), d_max)))
    return out

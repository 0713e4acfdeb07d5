"""Synthetic workload generators (CPU only): the seeded inputs of the batched,
band and serving measurements."""
import math
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pairwise-square-sums_amd"))
from ps4 import workloads  # noqa: E402


def test_splitmix64_known_values():
    # published first outputs of splitmix64 from seed 0 and seed 1234567
    g = workloads.splitmix64(0)
    assert [next(g) for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    g = workloads.splitmix64(1234567)
    assert [next(g) for _ in range(2)] == [6457827717110365317, 3203168211198807973]


def test_batched_ranges_shape():
    r = workloads.batched_ranges(4096, seed=42, top=10**11)
    assert len(r) == 4096
    assert all(1 <= a < b <= 10**11 + 1 for a, b in r)
    lens = [b - a for a, b in r]
    assert min(lens) >= 10**4 and max(lens) <= 10**9
    # log-uniform: the median length is near 10^6.5
    lens.sort()
    assert 10**6 < lens[2048] < 10**7
    assert r == workloads.batched_ranges(4096, seed=42, top=10**11)   # deterministic


def test_band_ranges_cover_exactly():
    sub = workloads.band_ranges(9 * 10**11, 10**12, 10**8)
    assert len(sub) == 1000 and sub[0][0] == 9 * 10**11 and sub[-1][1] == 10**12
    assert all(x[1] == y[0] for x, y in zip(sub, sub[1:]))


def test_serving_stream_requests():
    s = workloads.serving_stream(3000, rate=200.0, seed=7)
    times = [t for t, _, _ in s]
    assert times == sorted(times)
    assert 13.0 < times[-1] < 17.0                        # 3000 arrivals at 200/s
    kinds = [k for _, k, _ in s]
    assert 0.65 < kinds.count("range") / len(kinds) < 0.75
    for _, kind, args in s:
        if kind == "range":
            lo, hi = args
            assert 1 <= lo < 10**12 and 10**3 <= hi - lo <= 10**8
        else:
            (a, b, c), d_max = args
            assert 0 < a < b < c <= 10**9 and d_max == 10**12
            for x, y in ((a, b), (a, c), (b, c)):
                assert math.isqrt(x + y) ** 2 == x + y

"""The allocation-free triple query (ps4_extend_triple_noalloc) against the
ordinary one (ps4_extend_triple), through the C ABI -- needs an MI355X (-m gpu).

Two calls in a row on one device, the second with a smaller input (a smaller
d_max, so fewer answers), each equal to the ordinary entry point's answer:
the fixed scratch keeps the previous call's contents, so this checks that a
call clears what must start at zero and reads only its own answers."""
import random
import sys

import pytest

import ora
import ps4

pytestmark = pytest.mark.gpu


def say(msg):
    sys.__stderr__.write(f"[gpu-test] {msg}\n")
    sys.__stderr__.flush()


def _triples(n, seed):
    _, lines = ora.census(10**9, 10**9 + 3000, noext=True, dump4=2**62)
    quads = [[int(x) for x in ln.split("[")[1].split("]")[0].split(",")] for ln in lines]
    rng = random.Random(seed)
    rng.shuffle(quads)
    out = []
    for v in quads[:n]:
        drop = rng.randrange(4)
        out.append([x for i, x in enumerate(v) if i != drop])
    return out


def test_noalloc_equals_ordinary_twice_in_a_row():
    e = ps4.Engine(S_hi_max=10**12 + 10**8)
    try:
        triples = _triples(12, 5)
        for abc in triples:
            say(f"noalloc {abc}")
            big = e.extend_triple_noalloc(abc, 10**12)
            assert big == e.extend_triple(abc, 10**12), abc
            small = e.extend_triple_noalloc(abc, 10**8)         # fewer answers after more
            assert small == e.extend_triple(abc, 10**8), abc
            assert small == [d for d in big if d <= 10**8]
        # a triple with answers, then one with none: no stale answers
        a = e.extend_triple_noalloc([482, 359, 2], 10**6)
        assert a == e.extend_triple([482, 359, 2], 10**6) and len(a) >= 1
        assert e.extend_triple_noalloc([482, 359, 2], 3) == e.extend_triple([482, 359, 2], 3) == []
    finally:
        e.close()


def test_noalloc_limits():
    e = ps4.Engine(S_hi_max=10**12 + 10**8)
    try:
        with pytest.raises(ps4.PS4Error) as ex:
            e.extend_triple_noalloc([1, 2, 3], 100)                   # not a square triple
        assert ex.value.rc == -1
        abc = _triples(40, 9)
        big = max(abc, key=lambda t: len(e.extend_triple(t, 10**12)))
        n = len(e.extend_triple(big, 10**12))
        if n >= 2:
            with pytest.raises(ps4.PS4Error) as ex:
                e.extend_triple_noalloc(big, 10**12, cap=n - 1)        # d_out too small
            assert ex.value.rc == -2
        assert e.extend_triple_noalloc(big, 10**12, cap=n) == e.extend_triple(big, 10**12)
    finally:
        e.close()

"""C-ABI surface and host-side constants -- CPU only, no device calls."""
import ctypes
import os
import math
import random
import re

import pytest

import ps4

REPO = ps4.REPO_DIR
HEADER = os.path.join(REPO, "include", "ps4.h")
KERNELS = os.path.join(ps4.PKG_DIR, "csrc", "ps4_kernels.hip")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(ps4_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("ps4_open", "ps4_search", "ps4_search_batch", "ps4_extend_triple", "ps4_close",
                 "ps4_strerror", "ps4_last_error"):
        assert must in names


def test_library_loads_and_exports_every_declared_symbol():
    lib = ctypes.CDLL(ps4.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_strerror_texts():
    lib = ps4.load()
    for rc in (0, -1, -2, -3, -4, -5, -6, -7):
        assert lib.ps4_strerror(rc)
    assert b"NOT certified" in lib.ps4_strerror(-2)


def test_cli_binary_present():
    assert os.access(ps4.CLI_PATH, os.X_OK)


def test_struct_layouts_match_header():
    # ps4_counters: 10 + 100 + 3 u64; ps4_config as declared
    assert ctypes.sizeof(ps4.Counters) == 113 * 8
    assert ps4.Config.S_hi_max.offset == 8 and ps4.Config.dump4_limit.offset == 24


def test_tuning_struct_layout():
    # ps4_tuning: eight u32 in the declared order
    assert ctypes.sizeof(ps4.Tuning) == 32
    assert ps4.Tuning.rq_block.offset == 8 and ps4.Tuning.n_cu.offset == 20 and ps4.Tuning.ext_chunk.offset == 28


def test_tuning_file_round_trip_and_errors():
    """The library's parser (ps4_parse_tuning) reads what format_tuning writes,
    ignores comments and blank lines, and rejects unknown keys and bad values."""
    d = {"rq_grid": 1024, "ext_grid": 2816, "rq_block": 512, "sw_shift": 1, "batch_pct": 60, "rq_tile": 8,
         "ext_chunk": 16}
    assert ps4.parse_tuning(ps4.format_tuning(d, "autotune\nsecond line")) == d
    got = ps4.parse_tuning("\n  # comment only\n ext_grid=12   # trailing\r\n\n")
    assert got == dict.fromkeys(ps4.Tuning.KEYS, 0) | {"ext_grid": 12}
    assert ps4.parse_tuning("") == dict.fromkeys(ps4.Tuning.KEYS, 0)
    for bad in ("foo = 1", "rq_grid = x", "rq_grid 5", "rq_grid =", "rq_grid = -1", "rq_grid = 4294967296",
                "rq_tile = 99999999999999999999", "lds_fac_above = 1", "= 3"):
        with pytest.raises(ps4.PS4Error):
            ps4.parse_tuning(bad)
    assert ps4.parse_tuning("ext_chunk = 1")["ext_chunk"] == 1


@pytest.mark.parametrize("name,m", [("mod11", 11), ("mod17", 17), ("mod19", 19), ("mod23", 23)])
def test_kernel_residue_constants_are_exact(name, m):
    """s mod 11, 17, 19, 23 by the 24-bit multiply-high in the extension
    kernel, for every s < 2^17 (its representatives of a class mod
    81719 = 11*17*19*23 are below 81719)."""
    import numpy as np
    src = open(KERNELS).read()
    mt = re.search(name + r"\(u32 s\) \{ return PS4_MAD_I24\(mulhi_u24\(s << 7, (\d+)u\), -(\d+), s\); \}", src)
    assert mt, name
    magic, mm = int(mt.group(1)), int(mt.group(2))
    assert mm == m and magic < 2**24
    s = np.arange(1 << 17, dtype=np.uint64)
    assert int(s[-1] << np.uint64(7)) < 2**24                         # the first factor fits 24 bits
    q = ((s << np.uint64(7)) * np.uint64(magic)) >> np.uint64(32)
    assert np.array_equal(s - np.uint64(m) * q, s % np.uint64(m))
    assert 11 * 17 * 19 * 23 == 81719 and "#define M2 81719.0" in src
    # the kernel clamps its representatives to the domains checked here, so no
    # residue can index past its table whatever the input
    assert "min(mod_m2(d1), 81718u)" in src and "min(mod_m2(d2), 81718u)" in src
    assert "rep_m1(d1) & 0xFFFFFu" in src and "rep_m1(d2) & 0xFFFFFu" in src


def test_folded_mask_mod_23_is_exact():
    """The late mask mod 23 keeps 12 bits, bit min(r, 23 - r) for u = r mod 23:
    u^2 + c is a square mod 23 exactly when (23 - u)^2 + c is."""
    sq = {r * r % 23 for r in range(23)}
    for c in range(23):
        full = [(r * r + c) % 23 in sq for r in range(23)]
        assert all(full[r] == full[min(r, 23 - r)] for r in range(23))
    assert 12 + 12 <= 25   # bits 12..23 of the late word, below the mask mod 17 at bit 25


def test_shifted_mask_mod_65_is_exact():
    """The candidate loop tests u + 64512, whose residue b = (u + 64512) mod 65
    is taken mod 64 as the mask's bit, and the host's table gives bit b the
    test of r = (b + 33) mod 65.  That is exact when r^2 = u^2 mod 65 for every
    u, and the offset leaves the residues mod 64 and 63 alone."""
    src = open(KERNELS).read()
    lib = open(os.path.join(ps4.PKG_DIR, "csrc", "ps4_lib.cpp")).read()
    off = int(float(re.search(r"#define EX_UOFF ([\d.]+)", src).group(1)))
    rot = int(re.search(r"const u32 r = \(b \+ (\d+)\) % 65; if \(isq\(\(r \* r \+ a\) % 65, 65\)\) qr\.t65\[a\]", lib).group(1))
    assert off % 64 == 0 and off % 63 == 0
    for u in range(65):
        b = (u + off) % 65 % 64
        r = (b + rot) % 65
        assert (r * r - u * u) % 65 == 0, u
    assert "bit64(q2.x, q2.y, r65 & 63u)" in src


def test_candidate_representative_mod_262080_is_in_range():
    """x - 262080 * floor(x * EX_INV_M1) lies in [0, 2 * 262080) for integers
    0 <= x < 2^53, with EX_INV_M1 = (1/262080) (1 - 2^-45) in doubles: it is
    below 1/262080 by more than the product's rounding, and by little enough
    that the floor is at most one below.  Checked at the bound's extremes:
    multiples of 262080 and their neighbours, up to 2^53."""
    from fractions import Fraction
    import numpy as np
    src = open(KERNELS).read()
    assert "#define EX_INV_M1 ((1.0 / 262080.0) * (1.0 - 0x1p-45))" in src
    m = 262080
    inv = (1.0 / 262080.0) * (1.0 - 2.0**-45)
    dev = 1 - Fraction(inv) * m                      # relative deficit of inv
    assert Fraction(2**-45) - Fraction(2**-52) <= dev <= Fraction(2**-45) + Fraction(2**-52)
    rng = random.Random(5)
    ks = [0, 1, 2, 3] + [rng.randrange(1, 2**53 // m) for _ in range(20000)] + [2**53 // m - j for j in range(1, 4)]
    xs = np.array(sorted({k * m + d for k in ks for d in (-2, -1, 0, 1, 2) if 0 <= k * m + d < 2**53}), dtype=np.int64)
    q = np.floor(xs.astype(np.float64) * inv).astype(np.int64)
    r = xs - q * m
    assert r.min() >= 0 and r.max() < 2 * m
    assert 2 * m < (1 << 20)   # the domain of the mod 63 and mod 65 multiply-highs


@pytest.mark.parametrize("name,m", [("mod63", 63), ("mod65", 65)])
def test_kernel_residues_mod_63_65_are_exact(name, m):
    """s mod 63 and s mod 65 by the 24-bit multiply-high, for every s < 2^20 (the
    representatives of a class mod 262080 that the extension kernel forms are
    below 3 * 262080 < 2^20 in the set-up and below 2 * 262080 in the
    candidate loop)."""
    import numpy as np
    src = open(KERNELS).read()
    mt = re.search(name + r"\(u32 s\) \{ return PS4_MAD_I24\(mulhi_u24\(s << 3, (\d+)u\), -(\d+), s\); \}", src)
    assert mt, name
    magic, mm = int(mt.group(1)), int(mt.group(2))
    assert mm == m and magic < 2**24
    s = np.arange(1 << 20, dtype=np.uint64)
    assert int(s[-1] << np.uint64(3)) < 2**24                         # the first factor fits 24 bits
    q = ((s << np.uint64(3)) * np.uint64(magic)) >> np.uint64(32)
    assert np.array_equal(s - np.uint64(m) * q, s % np.uint64(m))
    assert 3 * 262080 < (1 << 20)


def test_codegen_has_no_signed_shift_of_unsigned_products(tmp_path):
    """Regression guard for a bring-up fault: the gfx950 code of the kernels must
    not shift a multiply result arithmetically (with __umul24 the compiler emitted
    v_mul_u32_u24 + v_ashrrev_i32 by 19 or 20, turning residues >= 2^31/magic into
    indices ~90000 and reading far past the residue tables).  An arithmetic shift
    by 31 is a sign mask (0 or all ones, whatever its operand: the compiler makes
    it from selects on the sign of a difference) and is allowed; any other
    amount is not."""
    import shutil
    import subprocess
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = tmp_path / "k.s"
    src = os.path.join(ps4.PKG_DIR, "csrc", "ps4_kernels.hip")
    subprocess.run([hipcc, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                    "-packed-fp32-ops", "-I" + os.path.join(ps4.PKG_DIR, "csrc"), "-I" + os.path.join(REPO, "include"),
                    "--cuda-device-only", "-S", src, "-o", str(out)], check=True, stdout=subprocess.PIPE,
                   stderr=subprocess.PIPE)
    asm = out.read_text()
    shifts = re.findall(r"v_ashrrev_i32(?:_e32|_e64)?\s+v\d+,\s*([^,\s]+),", asm)
    assert len(shifts) == asm.count("v_ashrrev_i32")   # every one parsed
    assert all(a == "31" for a in shifts), sorted(set(shifts))
    assert "__umul24(" not in open(src).read()


def test_division_magic_of_the_candidate_loop_is_exact():
    """t / nL = ((t << 9) * ceil(2^23 / nL)) >> 32 for every t < 8192 and
    1 <= nL < 1024 (the extension kernel's index of a candidate in its lists)."""
    import numpy as np
    t = np.arange(8192, dtype=np.uint64)
    for d in range(1, 1024):
        m = (0x800000 + d - 1) // d
        assert m < (1 << 24)
        assert np.array_equal(((t << np.uint64(9)) * np.uint64(m)) >> np.uint64(32), t // np.uint64(d)), d
    assert (8191 << 9) < (1 << 24)   # both factors fit the 24-bit multiply


def test_list_offsets_of_the_candidate_loop_are_exact():
    """8 (t mod nL) is the low half of j * w + 8 t for j = t / nL and any word w
    whose low half is -8 nL mod 2^16 (the record keeps the candidate offset in
    the high half)."""
    import numpy as np
    t = np.arange(8192, dtype=np.uint64)
    for d in range(1, 1024):
        w = np.uint64(((-8 * d) & 0xFFFF) | (8191 << 16))
        j = t // np.uint64(d)
        got = (j * w + (t << np.uint64(3))) & np.uint64(0xFFFF)
        assert np.array_equal(got, np.uint64(8) * (t % np.uint64(d))), d


def test_divisor_list_arena_holds_every_4set():
    """The largest nL + nH of one 4-set fits the extension kernel's arena: over
    every factor sequence of D' (the power of 2 first, radix <= 49, then odd
    primes in any order, radix <= 33) with d(D') <= 16384, the best prefix split
    (first minimum of nL + nH, as the kernel chooses it) needs at most 748."""
    src = open(os.path.join(REPO, "pairwise-square-sums_amd", "csrc", "ps4_kernels.hip")).read()
    arena = int(re.search(r"#define EX_ARENA (\d+)", src).group(1))
    worst = [0]

    def split(seq, n):
        best, nl = 1 + n, 1
        for r in seq:
            nl *= r
            best = min(best, nl + n // nl)
        return best

    def rec(seq, prod):
        if seq:
            worst[0] = max(worst[0], split(seq, prod))
        for r in range(2, (49 if not seq else 33) + 1):
            if prod * r > 16384:
                break
            seq.append(r)
            rec(seq, prod * r)
            seq.pop()

    rec([], 1)
    assert worst[0] == 748
    assert worst[0] <= arena


def test_triple_count_of_a_bucket_is_exact_in_32_bits():
    """C(k,3) = mulhi(C(k,2) * (k - 2), 0xAAAAAAAB) >> 1 with every step in 32
    bits for 3 <= k <= 2000 (the representation kernel's triple count of a
    bucket; larger buckets are refused before it is used)."""
    import math
    for k in range(3, 2001):
        x = (k * (k - 1) >> 1) * (k - 2)
        assert x < 1 << 32
        assert ((x * 0xAAAAAAAB) >> 32) >> 1 == math.comb(k, 3), k


def _walk_run(X0, X1, p_first, p_last):
    """The representation kernel's column walk over columns p_first..p_last of
    the superwindow [X0, X1), step for step in 32-bit arithmetic: the start from
    an integer square root, then per column the points at q, q+1, ... while
    p^2 + q^2 < X1, and the move to the next column with at most one step of q
    down (or onto the diagonal).  Returns the points (p, q)."""
    M = 0xFFFFFFFF
    ulen = X1 - X0
    pp = p_first * p_first
    if X0 <= 2 * pp:
        q = p_first
    else:
        t = X0 - pp
        q = math.isqrt(t)
        if q * q < t:
            q += 1
    o = pp + q * q - X0
    assert 0 <= o < 1 << 32
    pc, pts = p_first, []
    while pc <= p_last:
        oo, qq = o, q
        while oo < ulen:
            pts.append((pc, qq))
            oo = (oo + 2 * qq + 1) & M
            qq += 1
        pn = pc + 1
        t = (2 * q - 1) & M
        o = (o + 2 * pc + 1) & M
        dec = (((pn - q) & M) & (~((o - t) & M) & M)) >> 31
        diag = ((q - pn) & M) >> 31
        o = (o - (t if dec else 0) + ((2 * pn - 1) if diag else 0)) & M
        q = q - dec + diag
        pc = pn
    return pts


def test_lattice_walk_matches_brute_force():
    """The column walk of the representation kernel finds exactly the lattice
    points 1 <= p <= q, X0 <= p^2 + q^2 < X1: whole superwindows at small S
    (every column, including the diagonal and X0 < 64), and runs of columns at
    high S (up to 1e13) against the direct least root of every column.  Pins
    the claim that one column step moves q down at most once."""
    import math
    rng = random.Random(7)
    cases = [(1, 64), (2, 3), (5, 4101), (4096, 8192), (60, 64)]
    cases += [(x, x + rng.choice([64, 4096, 1 << 14])) for x in (rng.randrange(1, 10**7) for _ in range(30))]
    for X0, X1 in cases:
        pmax = math.isqrt((X1 - 1) // 2)
        got = _walk_run(X0, X1, 1, pmax) if pmax >= 1 else []
        want = [(p, q) for p in range(1, pmax + 1)
                for q in range(max(p, math.isqrt(max(X0 - p * p, 0))), math.isqrt(X1 - 1 - p * p) + 1)
                if X0 <= p * p + q * q < X1]
        assert got == want, (X0, X1)
    for _ in range(40):
        X0 = rng.randrange(10**9, 10**13)
        X1 = X0 + rng.choice([4096, 1 << 17, 1 << 21])
        pmax = math.isqrt((X1 - 1) // 2)
        a = rng.randrange(1, pmax)
        b = min(pmax, a + 400)
        got = _walk_run(X0, X1, a, b)
        want = []
        for p in range(a, b + 1):
            q0 = max(p, math.isqrt(X0 - p * p - 1) + 1 if X0 > p * p else p)
            q = q0
            while p * p + q * q < X1:
                if p * p + q * q >= X0:
                    want.append((p, q))
                q += 1
        assert got == want, (X0, X1, a)


def test_divisor_count_correction_is_exact_in_single_precision():
    """When ra - rb and ra + rb share a prime, the extension kernel corrects
    d = d(o1) d(o2) by d / ((ea+1)(eb+1)) * (ea+eb+1) with a single-precision
    quotient: rint(d * rcp(den)).  Exact when den | d, d <= 2^16 (no odd
    n < 2^26 + 64 has more than 256 divisors) and the reciprocal and the
    product are within 2^-22 relative of the true values (1 ulp each)."""
    import numpy as np
    lim = (1 << 26) + 64
    primes = [3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47]
    best = [1]

    def dfs(i, n, d, emax):
        best[0] = max(best[0], d)
        if i == len(primes):
            return
        m, e = n, 0
        while e < emax and m * primes[i] < lim:
            m *= primes[i]
            e += 1
            dfs(i + 1, m, d * (e + 1), e)

    dfs(0, 1, 1, 64)
    assert best[0] == 256
    emax = 16   # 3^17 > 2^26 + 64
    for ea in range(emax + 1):
        for eb in range(emax + 1):
            den = (ea + 1) * (eb + 1)
            q = np.arange(1, (1 << 16) // den + 1, dtype=np.float64)
            d = q * den
            for rel in (-2.0**-22, 2.0**-22):
                approx = np.rint(d * (1.0 / den) * (1.0 + rel))
                assert np.array_equal(approx, q), (ea, eb)

"""ctypes binding of the CPU restatement (oracle/_bin/libps4oracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, always as the checker.  Build with
`make -C oracle ora`.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_bin", "libps4oracle.so")
CLI = os.path.join(HERE, "_bin", "ps4_oracle")
REF_S4 = os.path.join(HERE, "_ref", "s4")
TOTAL_KEYS = ["S_done", "pairs", "S_ge3", "triples", "4-sets", "5-sets", "6-sets",
              "multi-ext-4sets", "ext_cand", "ext_sqrt"]
_FIELDS = ["S_done", "pairs", "S_ge3", "triples", "sets4", "sets5", "sets6", "multi", "ext_cand", "ext_sqrt"]


class OraCounters(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_uint64) for n in _FIELDS] + [("hist5", ctypes.c_uint64 * 100)] +
                [(n, ctypes.c_uint64) for n in ("max_reps", "max_divisors", "max_extensions")])


LINE = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_char_p)
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: make -C oracle ora")
        lib = ctypes.CDLL(LIB)
        u64 = ctypes.c_uint64
        lib.ora_census.argtypes = [u64, u64, ctypes.c_uint, ctypes.c_int64, LINE, ctypes.c_void_p,
                                   ctypes.POINTER(OraCounters)]
        lib.ora_extend_triple.argtypes = [ctypes.POINTER(u64), u64, ctypes.POINTER(u64), ctypes.c_size_t]
        lib.ora_extend_triple.restype = ctypes.c_int64
        lib.ora_isqrt64.argtypes = [u64]
        lib.ora_isqrt64.restype = u64
        lib.ora_isqrt128.argtypes = [u64, u64]
        lib.ora_isqrt128.restype = u64
        lib.ora_is_square128.argtypes = [u64, u64]
        lib.ora_num_divisors.argtypes = [u64]
        lib.ora_num_divisors.restype = ctypes.c_uint32
        lib.ora_divisors_of_product.argtypes = [u64, u64, ctypes.POINTER(u64), ctypes.c_int]
        lib.ora_square_triples.argtypes = [u64]
        lib.ora_square_triples.restype = u64
        lib.ora_elements.argtypes = [u64, LINE, ctypes.c_void_p]
        lib.ora_elements.restype = ctypes.c_int64
        _lib = lib
    return _lib


def census(lo, hi, noext=False, dump4=0):
    """(result dict like the reference's stderr report, sorted lines); rc != 0 -> RuntimeError."""
    lib = load()
    lines = []

    def cb(_u, line):
        lines.append(line.decode().rstrip("\n"))
    f = LINE(cb)
    c = OraCounters()
    rc = lib.ora_census(int(lo), int(hi), 1 if noext else 0, int(dump4), f, None, ctypes.byref(c))
    if rc:
        raise RuntimeError(f"oracle census failed rc={rc}")
    res = {"counters": {k: int(getattr(c, f_)) for k, f_ in zip(TOTAL_KEYS, _FIELDS)},
           "hist5": {str(i): int(c.hist5[i]) for i in range(100) if c.hist5[i]},
           "capacity": {"reps": int(c.max_reps), "divisors": int(c.max_divisors),
                        "extensions": int(c.max_extensions)}}
    return res, sorted(lines)


def extend_triple(abc, d_max, cap=1 << 16):
    lib = load()
    arr = (ctypes.c_uint64 * 3)(*[int(x) for x in abc])
    out = (ctypes.c_uint64 * cap)()
    n = lib.ora_extend_triple(arr, int(d_max), out, cap)
    if n < 0:
        raise ValueError("not a square triple")
    return [int(out[i]) for i in range(min(n, cap))]


def isqrt64(n):
    return int(load().ora_isqrt64(int(n)))


def isqrt128(n):
    return int(load().ora_isqrt128(n & (2**64 - 1), n >> 64))


def is_square128(n):
    return bool(load().ora_is_square128(n & (2**64 - 1), n >> 64))


def num_divisors(n):
    return int(load().ora_num_divisors(int(n)))


def divisors_of_product(m1, m2, cap=1 << 16):
    out = (ctypes.c_uint64 * cap)()
    n = load().ora_divisors_of_product(int(m1), int(m2), out, cap)
    if n < 0:
        raise ValueError("too many divisors")
    return [int(out[i]) for i in range(n)]


def square_triples(n_max):
    """Number of square triples {a<b<c} with c <= n_max (brute force)."""
    return int(load().ora_square_triples(int(n_max)))


def elements(n_max):
    """Every 4-set with largest element <= n_max, as sorted "[a,b,c,d]" lines (brute force)."""
    lib = load()
    lines = []

    def cb(_u, line):
        lines.append(line.decode().rstrip("\n"))
    f = LINE(cb)
    n = lib.ora_elements(int(n_max), f, None)
    if n < 0:
        raise RuntimeError("oracle element search failed")
    return lines

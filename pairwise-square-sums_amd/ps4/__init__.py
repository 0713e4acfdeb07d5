"""Python face of the MI355X census engine (ctypes over lib/libps4.so).

The engine is the reference's s4 census (exhaustive/s4.c of
panjanek/pairwise-square-sums) on one gfx950 device.  This module only marshals
arguments: every count and every result line comes from the HIP kernels behind
include/ps4.h.  If the shared library is missing the import of the binding
raises; there is no CPU fallback.

    from ps4 import Engine
    with Engine(S_hi_max=10**7) as e:
        counters, lines = e.search(1, 10**7)
"""
import ctypes
import os
import subprocess

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libps4.so")
CLI_PATH = os.path.join(PKG_DIR, "lib", "ps4")
REPO_DIR = os.path.dirname(PKG_DIR)

NOEXT, VERIFY, TEST6, GRAPH = 0x1, 0x2, 0x4, 0x8
ERRORS = {0: "PS4_OK", -1: "PS4_E_ARG", -2: "PS4_E_CAPACITY", -3: "PS4_E_HIP", -4: "PS4_E_OOM",
          -5: "PS4_E_VERIFY", -6: "PS4_E_SINK", -7: "PS4_E_INTERNAL"}
# TOTAL-line keys of the reference (exhaustive/s4.c:498-500), in order
TOTAL_KEYS = ["S_done", "pairs", "S_ge3", "triples", "4-sets", "5-sets", "6-sets",
              "multi-ext-4sets", "ext_cand", "ext_sqrt"]
_FIELDS = ["S_done", "pairs", "S_ge3", "triples", "sets4", "sets5", "sets6", "multi", "ext_cand", "ext_sqrt"]


class Counters(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_uint64) for n in _FIELDS] + [("hist5", ctypes.c_uint64 * 100)] +
                [(n, ctypes.c_uint64) for n in ("max_reps", "max_divisors", "max_extensions")])

    def as_dict(self):
        """TOTAL counters, 5-set histogram and capacity maxima, keyed like the
        reference's stderr report."""
        tot = {k: int(getattr(self, f)) for k, f in zip(TOTAL_KEYS, _FIELDS)}
        hist = {str(i): int(self.hist5[i]) for i in range(100) if self.hist5[i]}
        cap = {"reps": int(self.max_reps), "divisors": int(self.max_divisors),
               "extensions": int(self.max_extensions)}
        return {"counters": tot, "hist5": hist, "capacity": cap}


class Config(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("S_hi_max", ctypes.c_uint64), ("flags", ctypes.c_uint32),
                ("dump4_limit", ctypes.c_int64), ("cap_reps", ctypes.c_uint32), ("cap_divs", ctypes.c_uint32),
                ("cap_ext", ctypes.c_uint32)]


class Tuning(ctypes.Structure):
    """Launch shape and batch plan (ps4_tuning): results never depend on it."""
    _fields_ = [("rq_grid", ctypes.c_uint32), ("ext_grid", ctypes.c_uint32), ("rq_block", ctypes.c_uint32),
                ("sw_shift", ctypes.c_uint32), ("batch_pct", ctypes.c_uint32), ("n_cu", ctypes.c_uint32),
                ("rq_tile", ctypes.c_uint32), ("ext_chunk", ctypes.c_uint32)]
    KEYS = ("rq_grid", "ext_grid", "rq_block", "sw_shift", "batch_pct", "rq_tile", "ext_chunk")

    def as_dict(self):
        d = {k: int(getattr(self, k)) for k in self.KEYS}
        if self.n_cu:
            d["n_cu"] = int(self.n_cu)
        return d

    @classmethod
    def from_dict(cls, d):
        d = {k: v for k, v in d.items() if k != "n_cu"}
        unknown = set(d) - set(cls.KEYS)
        if unknown:
            raise ValueError(f"unknown tuning keys {sorted(unknown)}")
        return cls(**{k: int(v) for k, v in d.items()})


def parse_tuning(text):
    """Parse a tuning file's text with the library's own parser (no device needed)."""
    lib = load()
    t = Tuning()
    rc = lib.ps4_parse_tuning(text.encode(), ctypes.byref(t))
    if rc:
        raise PS4Error(rc, "tuning file")
    return t.as_dict()


def format_tuning(d, comment=""):
    """The text of a tuning file for the dict d (the grammar ps4_parse_tuning reads)."""
    head = "".join(f"# {c}\n" for c in comment.splitlines() if c)
    return head + "".join(f"{k} = {int(d[k])}\n" for k in Tuning.KEYS if k in d)


class Done(ctypes.Structure):
    """One finished request of the serving stream (ps4_done)."""
    _fields_ = [("tag", ctypes.c_uint64), ("rc", ctypes.c_int), ("kind", ctypes.c_int), ("counters", Counters),
                ("lines", ctypes.c_void_p), ("nbytes", ctypes.c_size_t), ("d", ctypes.POINTER(ctypes.c_uint64)),
                ("n_d", ctypes.c_size_t)]


class StreamState(ctypes.Structure):
    _fields_ = [("pending_ranges", ctypes.c_uint64), ("pending_triples", ctypes.c_uint64),
                ("quad_slots", ctypes.c_uint64), ("hit_slots", ctypes.c_uint64), ("triple_slots", ctypes.c_uint64)]


SINK = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t)
DONE = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(Done))
ALERT = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_char_p)

_lib = None


class PS4Error(RuntimeError):
    def __init__(self, rc, detail=""):
        self.rc = rc
        self.name = ERRORS.get(rc, str(rc))
        super().__init__(f"{self.name}: {detail}" if detail else self.name)


def load():
    """Load lib/libps4.so (built by __graft_entry__.build() / `make -C pairwise-square-sums_amd`)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build the HIP extension first")
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    u64, u64p = ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)
    lib.ps4_open.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(P)]
    lib.ps4_search.argtypes = [P, u64, u64, SINK, P, ctypes.POINTER(Counters)]
    lib.ps4_search_batch.argtypes = [P, ctypes.c_size_t, u64p, u64p, SINK, ctypes.POINTER(P),
                                     ctypes.POINTER(Counters)]
    lib.ps4_extend_triple.argtypes = [P, u64p, u64, u64p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    lib.ps4_extend_triple_noalloc.argtypes = [P, u64p, u64, u64p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    lib.ps4_extend_triple_batch.argtypes = [P, ctypes.c_size_t, u64p, u64, u64p, ctypes.c_size_t,
                                            ctypes.POINTER(ctypes.c_size_t)]
    lib.ps4_search_elements.argtypes = [P, u64, SINK, P, u64p, u64p]
    lib.ps4_submit_range.argtypes = [P, u64, u64, u64]
    lib.ps4_submit_triple.argtypes = [P, u64, u64p, u64]
    lib.ps4_step.argtypes = [P, u64, DONE, P, ctypes.POINTER(StreamState)]
    lib.ps4_set_alert.argtypes = [P, ALERT, P]
    lib.ps4_set_dump4.argtypes = [P, ctypes.c_int64]
    lib.ps4_get_tuning.argtypes = [P, ctypes.POINTER(Tuning), ctypes.POINTER(Tuning)]
    lib.ps4_set_tuning.argtypes = [P, ctypes.POINTER(Tuning)]
    lib.ps4_parse_tuning.argtypes = [ctypes.c_char_p, ctypes.POINTER(Tuning)]
    lib.ps4_close.argtypes = [P]
    lib.ps4_close.restype = None
    lib.ps4_strerror.argtypes = [ctypes.c_int]
    lib.ps4_strerror.restype = ctypes.c_char_p
    lib.ps4_last_error.argtypes = [P]
    lib.ps4_last_error.restype = ctypes.c_char_p
    lib.ps4_redo_count.argtypes = [P]
    lib.ps4_redo_count.restype = ctypes.c_uint64
    lib.ps4_kernel_time.argtypes = [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                    u64p, ctypes.c_int]
    _lib = lib
    return lib


class Engine:
    """One device context (resident tables for S < S_hi_max)."""

    def __init__(self, S_hi_max, device=0, noext=False, verify=False, test6=False, dump4=0,
                 cap_reps=0, cap_divs=0, cap_ext=0, tune=None, graph=False):
        """tune: a tuning dict or the path of a tuning file (ps4/autotune.py);
        graph: launch each batch as one captured hipGraph (PS4_GRAPH)."""
        self.lib = load()
        cfg = Config(device, int(S_hi_max), (NOEXT if noext else 0) | (VERIFY if verify else 0) |
                     (TEST6 if test6 else 0) | (GRAPH if graph else 0), int(dump4), cap_reps, cap_divs,
                     cap_ext)
        self.ctx = ctypes.c_void_p()
        rc = self.lib.ps4_open(ctypes.byref(cfg), ctypes.byref(self.ctx))
        if rc:
            raise PS4Error(rc, "ps4_open")
        self.alerts = []
        self._alert_cb = None
        if tune is not None:
            self.set_tuning(tune)

    def _check(self, rc):
        if rc:
            raise PS4Error(rc, self.lib.ps4_last_error(self.ctx).decode())

    def set_alert(self, collect=True):
        """Route 6-set / drill lines to self.alerts instead of the result lines."""
        if collect:
            def cb(_user, line):
                self.alerts.append(line.decode())
            self._alert_cb = ALERT(cb)
        else:
            self._alert_cb = ALERT(0)
        self._check(self.lib.ps4_set_alert(self.ctx, self._alert_cb, None))

    def get_tuning(self):
        """(values in force, maxima) as dicts."""
        cur, mx = Tuning(), Tuning()
        self._check(self.lib.ps4_get_tuning(self.ctx, ctypes.byref(cur), ctypes.byref(mx)))
        return cur.as_dict(), mx.as_dict()

    def set_tuning(self, tune):
        """Apply a tuning dict (missing or 0 fields: defaults) or a tuning file's path."""
        if isinstance(tune, (str, os.PathLike)):
            with open(tune) as f:
                tune = parse_tuning(f.read())
        t = Tuning.from_dict(tune)
        self._check(self.lib.ps4_set_tuning(self.ctx, ctypes.byref(t)))

    def set_dump4(self, n):
        self._check(self.lib.ps4_set_dump4(self.ctx, int(n)))

    def search(self, lo, hi):
        """Census of S in [lo, hi): (counter dict, list of result lines)."""
        lines = []

        # each call holds whole lines; the library makes a batch's lines while the
        # next batch runs, so splitting them here overlaps the device work
        def sink(_user, text, n):
            lines.extend(ctypes.string_at(text, n).decode().splitlines())
            return 0
        cb = SINK(sink)
        c = Counters()
        self._check(self.lib.ps4_search(self.ctx, int(lo), int(hi), cb, None, ctypes.byref(c)))
        return c.as_dict(), lines

    def search_batch(self, los, his):
        """Independent intervals in one persistent launch: list of (counters, lines)."""
        n = len(los)
        lo = (ctypes.c_uint64 * n)(*[int(x) for x in los])
        hi = (ctypes.c_uint64 * n)(*[int(x) for x in his])
        texts = [[] for _ in range(n)]
        users = (ctypes.c_void_p * n)(*range(1, n + 1))

        def sink(user, text, nb):   # whole lines per call (as in search)
            texts[int(user) - 1].extend(ctypes.string_at(text, nb).decode().splitlines())
            return 0
        cb = SINK(sink)
        cs = (Counters * n)()
        self._check(self.lib.ps4_search_batch(self.ctx, n, lo, hi, cb, users, cs))
        return [(cs[i].as_dict(), texts[i]) for i in range(n)]

    def search_elements(self, n_max):
        """Element-range search: (sorted "[a,b,c,d]" lines of every 4-set with
        largest element <= n_max, number of square triples with c <= n_max)."""
        chunks = []

        def sink(_user, text, n):
            chunks.append(ctypes.string_at(text, n))
            return 0
        cb = SINK(sink)
        ns, nt = ctypes.c_uint64(), ctypes.c_uint64()
        self._check(self.lib.ps4_search_elements(self.ctx, int(n_max), cb, None, ctypes.byref(ns), ctypes.byref(nt)))
        lines = b"".join(chunks).decode().splitlines()
        if len(lines) != ns.value:
            raise RuntimeError("element search: line count differs from the set count")
        return lines, int(nt.value)

    def extend_triple(self, abc, d_max, cap=1 << 16):
        arr = (ctypes.c_uint64 * 3)(*[int(x) for x in abc])
        out = (ctypes.c_uint64 * cap)()
        n = ctypes.c_size_t()
        self._check(self.lib.ps4_extend_triple(self.ctx, arr, int(d_max), out, cap, ctypes.byref(n)))
        return [int(out[i]) for i in range(min(n.value, cap))]

    def extend_triple_noalloc(self, abc, d_max, cap=1 << 16):
        """The same query through ps4_extend_triple_noalloc (no device
        allocation: the library's fixed per-device scratch)."""
        arr = (ctypes.c_uint64 * 3)(*[int(x) for x in abc])
        out = (ctypes.c_uint64 * cap)()
        n = ctypes.c_size_t()
        self._check(self.lib.ps4_extend_triple_noalloc(self.ctx, arr, int(d_max), out, cap, ctypes.byref(n)))
        return [int(out[i]) for i in range(n.value)]

    def extend_triple_batch(self, triples, d_max, cap=4096):
        """Triple queries in one launch: one ascending list of d per triple (a
        query with more than cap answers is asked again alone, in full)."""
        n = len(triples)
        arr = (ctypes.c_uint64 * (3 * n))(*[int(x) for t in triples for x in t])
        out = (ctypes.c_uint64 * (n * cap))()
        cnt = (ctypes.c_size_t * n)()
        self._check(self.lib.ps4_extend_triple_batch(self.ctx, n, arr, int(d_max), out, cap, cnt))
        return [[int(out[i * cap + k]) for k in range(cnt[i])] if cnt[i] <= cap
                else self.extend_triple(triples[i], d_max) for i in range(n)]

    # ---- serving stream (ps4_submit_*, ps4_step): the caller drives the steps
    def submit_range(self, tag, lo, hi):
        self._check(self.lib.ps4_submit_range(self.ctx, int(tag), int(lo), int(hi)))

    def submit_triple(self, tag, abc, d_max):
        arr = (ctypes.c_uint64 * 3)(*[int(x) for x in abc])
        self._check(self.lib.ps4_submit_triple(self.ctx, int(tag), arr, int(d_max)))

    def step(self, max_quads=0):
        """One step of the stream: a list of (tag, error or None, result) for
        the requests it finished (result: (counters, lines) or the list of d),
        and the state after it (dict)."""
        out = []

        def cb(_user, dp):
            d = dp.contents
            if d.rc:
                out.append((int(d.tag), PS4Error(d.rc, self.lib.ps4_last_error(self.ctx).decode()), None))
            elif d.kind == 0:
                text = ctypes.string_at(d.lines, d.nbytes).decode() if d.nbytes else ""
                out.append((int(d.tag), None, (d.counters.as_dict(), text.splitlines())))
            else:
                out.append((int(d.tag), None, [int(d.d[i]) for i in range(d.n_d)]))
        f = DONE(cb)
        st = StreamState()
        self._check(self.lib.ps4_step(self.ctx, int(max_quads), f, None, ctypes.byref(st)))
        return out, {k: int(getattr(st, k)) for k, _ in StreamState._fields_}

    def redo_count(self):
        """Batches redone after a device queue overflowed, since the engine opened."""
        return int(self.lib.ps4_redo_count(self.ctx))

    def kernel_time(self, reset=False):
        """(ms in k_rq, ms in k_ext, batches) accumulated on the device stream."""
        a, b, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_uint64()
        self._check(self.lib.ps4_kernel_time(self.ctx, ctypes.byref(a), ctypes.byref(b), ctypes.byref(n),
                                             1 if reset else 0))
        return a.value, b.value, n.value

    def close(self):
        if getattr(self, "ctx", None) and self.ctx.value:
            self.lib.ps4_close(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_cli(args, cwd=None, timeout=600):
    """Run the drop-in CLI (lib/ps4) with s4 arguments; returns CompletedProcess."""
    if not os.path.exists(CLI_PATH):
        raise RuntimeError(f"{CLI_PATH} is missing: build first")
    return subprocess.run([CLI_PATH] + [str(a) for a in args], cwd=cwd, stdout=subprocess.PIPE,
                          stderr=subprocess.PIPE, timeout=timeout)


def build(arch="gfx950", jobs=8):
    """Compile lib/libps4.so and lib/ps4 for the given gfx target (always rebuilds)."""
    subprocess.run(["make", "-B", f"-j{jobs}", "-C", PKG_DIR, f"ARCH={arch}"], check=True,
                   stdout=subprocess.PIPE, stderr=subprocess.STDOUT)

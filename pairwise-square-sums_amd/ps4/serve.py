"""A request loop over one census engine: range queries and triple queries as
they arrive, with continuous batching.

The reference answers one interval per process run (exhaustive/s4.c:416-426);
this loop keeps one device context open (factor tables, queues and output
buffers stay resident) and serves many small requests:

  * submit() enqueues a request and returns at once; a worker thread runs
    steps.  A step takes every queued request, oldest first, as long as the
    step's reservation of device output memory stays within the budget, runs
    all admitted range queries as one ps4_search_batch (their superwindows
    share the persistent launches, so their work overlaps on the device) and
    all admitted triple queries as one ps4_extend_triple_batch, then completes
    them.  Requests arriving meanwhile wait for the next step.
  * Device output memory per request: a range query reserves room for its
    expected 4-sets in the 4-set queue (32 B each, from an upper bound of the
    4-set density near its S) and a triple query its slot of the triple
    output buffer (65536 values of 8 B).  A request larger than the whole
    budget runs alone.
  * Results are exactly those of the same request run alone (each interval
    keeps its own counters and lines; triple answers are per query).
"""
import math
import threading
import time

TRIPLE_SLOT_BYTES = (1 << 16) * 8
QUAD_BYTES = 32


def quad_density(S):
    """Upper bound on 4-sets per S near S (the library sizes its batches by the same bound)."""
    l2 = math.log2(max(2, S))
    return min(0.7, max(0.08, 0.02 * l2 - 0.3))


class Request:
    __slots__ = ("kind", "args", "t_arrive", "t_admit", "t_done", "result", "error", "_ev")

    def __init__(self, kind, args, t_arrive):
        self.kind, self.args, self.t_arrive = kind, args, t_arrive
        self.t_admit = self.t_done = None
        self.result = self.error = None
        self._ev = threading.Event()

    def wait(self, timeout=None):
        return self._ev.wait(timeout)

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_arrive

    def reserve_bytes(self):
        if self.kind == "range":
            lo, hi = self.args
            return int(quad_density(hi) * (hi - lo) * QUAD_BYTES) + QUAD_BYTES
        return TRIPLE_SLOT_BYTES


class Server:
    """Continuous-batching loop over one ps4.Engine (the engine must not be
    used by anyone else while the server runs)."""

    def __init__(self, engine, out_budget_bytes=256 << 20, clock=time.perf_counter):
        self.eng = engine
        self.budget = out_budget_bytes
        self.clock = clock
        self._q = []
        self._cv = threading.Condition()
        self._stop = False
        self.steps = []          # (n_range, n_triple, seconds) per step
        self._th = threading.Thread(target=self._loop, name="ps4-serve", daemon=True)
        self._th.start()

    def submit(self, kind, args, t_arrive=None):
        if kind not in ("range", "triple"):
            raise ValueError(kind)
        r = Request(kind, args, self.clock() if t_arrive is None else t_arrive)
        with self._cv:
            self._q.append(r)
            self._cv.notify()
        return r

    def close(self):
        with self._cv:
            self._stop = True
            self._cv.notify()
        self._th.join()

    def _admit(self):
        """Oldest first while the reservations fit the budget (at least one)."""
        step, used = [], 0
        while self._q:
            need = self._q[0].reserve_bytes()
            if step and used + need > self.budget:
                break
            step.append(self._q.pop(0))
            used += need
        return step

    def _loop(self):
        while True:
            with self._cv:
                while not self._q and not self._stop:
                    self._cv.wait()
                if self._stop and not self._q:
                    return
                step = self._admit()
            t0 = self.clock()
            for r in step:
                r.t_admit = t0
            rng = [r for r in step if r.kind == "range"]
            tri = [r for r in step if r.kind == "triple"]
            try:
                if rng:
                    res = self.eng.search_batch([r.args[0] for r in rng], [r.args[1] for r in rng])
                    for r, x in zip(rng, res):
                        r.result = x
                if tri:
                    by_dmax = {}
                    for r in tri:
                        by_dmax.setdefault(r.args[1], []).append(r)
                    for d_max, rs in by_dmax.items():
                        res = self.eng.extend_triple_batch([r.args[0] for r in rs], d_max)
                        for r, x in zip(rs, res):
                            r.result = x
            except Exception as e:   # every request of the failed step reports it
                for r in step:
                    r.error = e
            t1 = self.clock()
            self.steps.append((len(rng), len(tri), t1 - t0))
            for r in step:
                r.t_done = t1
                r._ev.set()

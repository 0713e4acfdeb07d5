"""A request loop over one census engine: range queries and triple queries as
they arrive, with continuous batching in the library's serving stream.

The reference answers one interval per process run (exhaustive/s4.c:416-426);
this loop keeps one device context open (factor tables, queues and output
buffers stay resident) and serves many requests at once:

  * submit() enqueues a request and returns at once; a worker thread hands
    every request that has arrived to the library (ps4_submit_range /
    ps4_submit_triple) and runs steps (ps4_step) while any is unfinished.
  * A step is one short device batch.  The range requests, shortest remaining
    first, each take the next part of their interval while the step's expected
    4-sets fit its budget and the free part of the 4-set queue (output memory
    on the device), so a large request is cut into parts and never holds the
    device for long; every part of every admitted request shares the step's
    persistent launches.  The pending triple queries run in the same step on a
    second stream.  Requests that arrive during a step join the next one.
  * A request's result is exactly that of the same request run alone: a range
    request's counters are summed over its parts (they are additive over any
    partition of the range) and its lines collected; a triple request is one
    query.  A request the library refuses (bad range, not a square triple)
    fails alone, with its error; so does one whose part fails on the device.
"""
import threading
import time


class Request:
    __slots__ = ("kind", "args", "t_arrive", "t_done", "result", "error", "_ev")

    def __init__(self, kind, args, t_arrive):
        self.kind, self.args, self.t_arrive = kind, args, t_arrive
        self.t_done = None
        self.result = self.error = None
        self._ev = threading.Event()

    def wait(self, timeout=None):
        return self._ev.wait(timeout)

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_arrive


class Server:
    """Continuous-batching loop over one ps4.Engine (the engine must not be
    used by anyone else while the server runs).  step_quads: the budget of
    expected 4-sets per step (0: the library's default, 2^26: a whole 10^8-wide
    request at 10^12)."""

    def __init__(self, engine, step_quads=0, clock=time.perf_counter):
        self.eng = engine
        self.step_quads = step_quads
        self.clock = clock
        self._new = []
        self._live = {}           # tag -> Request handed to the library
        self._tag = 0
        self._cv = threading.Condition()
        self._stop = False
        self.steps = []           # (requests finished, seconds) per step
        self.capacity = None      # the library's per-step capacities (last step)
        self._th = threading.Thread(target=self._loop, name="ps4-serve", daemon=True)
        self._th.start()

    def submit(self, kind, args, t_arrive=None):
        if kind not in ("range", "triple"):
            raise ValueError(kind)
        r = Request(kind, args, self.clock() if t_arrive is None else t_arrive)
        with self._cv:
            self._new.append(r)
            self._cv.notify()
        return r

    def close(self):
        with self._cv:
            self._stop = True
            self._cv.notify()
        self._th.join()

    def _finish(self, r, result=None, error=None):
        r.result, r.error = result, error
        r.t_done = self.clock()
        r._ev.set()

    def _loop(self):
        import ps4
        while True:
            with self._cv:
                while not self._new and not self._live and not self._stop:
                    self._cv.wait()
                if self._stop and not self._new and not self._live:
                    return
                new, self._new = self._new, []
            for r in new:              # into the library's stream (refusals fail alone)
                tag = self._tag
                self._tag += 1
                try:
                    if r.kind == "range":
                        self.eng.submit_range(tag, *r.args)
                    else:
                        self.eng.submit_triple(tag, *r.args)
                except ps4.PS4Error as e:
                    self._finish(r, error=e)
                    continue
                self._live[tag] = r
            if not self._live:
                continue
            t0 = self.clock()
            try:
                done, state = self.eng.step(self.step_quads)
            except Exception as e:      # the step itself failed: every live request reports it
                for r in self._live.values():
                    self._finish(r, error=e)
                self._live.clear()
                continue
            self.capacity = state
            for tag, err, res in done:
                self._finish(self._live.pop(tag), result=res, error=err)
            self.steps.append((len(done), self.clock() - t0))

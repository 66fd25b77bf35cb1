"""Served solves: a request loop with continuous batching and HBM-budget admission.

Requests (an urban tile and a storm) arrive over time. Each admitted request becomes a slot of one
device-resident `Pool`; every loop iteration advances all active slots together by one chunk of
their own simulated time (one launch sequence for all of them), so requests that arrive while
others run join the next iteration as soon as the budget has room, and leave the batch as they
finish. A request's chunk schedule and rain are its own (rain held per chunk at the hyetograph's
mid-chunk value, as the reference app holds forcing per 10 s frame, app/map.js:2281), so its
result does not depend on what else is in the batch.

The reference has no serving path (src/batch.js runs one job after another); this module and its
test are the build's own.
"""
import collections
import time

import numpy as np

from . import scenarios as sc
from .solver import Pool


class Request:
    def __init__(self, rid, arrival, layers, duration, rain_fn, meta=None):
        self.id = rid
        self.arrival = arrival          # s, on the server's clock
        self.layers = layers            # nx, ny, dx, z, mask, manning, bc
        self.duration = duration        # simulated s
        self.rain_fn = rain_fn          # t (s) -> m/s
        self.meta = meta or {}
        self.slot = None
        self.t = 0.0
        self.admitted = None
        self.done = None
        self.result = None

    @property
    def cells(self):
        return self.layers["nx"] * self.layers["ny"]


def storm_request(rid, arrival, side, seed, depth_scale, duration=1800.0, dx=1.0, storm_s=1800.0):
    """A served solve of the urban generator at `side` under the peaked 30-minute storm
    (scenarios.peaked_hyetograph: 120 mm/h at 40 %, 25 mm), its rates scaled by depth_scale; the solve covers
    the first `duration` seconds of the `storm_s` storm (all of it by default)."""
    u = sc.urban(side, seed=seed, dx=dx)
    layers = dict(nx=u["nx"], ny=u["ny"], dx=dx, z=u["z"], mask=u["mask"], manning=u["manning"],
                  bc={k: "outflow" for k in ("west", "east", "south", "north")})
    rain = lambda t: depth_scale * float(sc.peaked_hyetograph(t, duration=storm_s))
    return Request(rid, arrival, layers, duration, rain, dict(side=side, seed=seed, depth_scale=depth_scale))


class Server:
    """Continuous batching over one Pool. `clock` gives the server's time in seconds (wall clock by
    default; tests pass a virtual clock)."""

    def __init__(self, budget_bytes, chunk_s=60.0, K=64, rows=64, clock=None, keep=("hmax",)):
        self.pool = Pool(budget_bytes)
        self.budget = budget_bytes
        self.chunk = float(chunk_s)
        self.K = K
        self.rows = rows
        self.clock = clock or time.perf_counter
        self.t0 = self.clock()
        self.waiting = collections.deque()
        self.active = []
        self.finished = []
        self.keep = keep
        self.iterations = 0
        self.cell_updates = 0

    def now(self):
        return self.clock() - self.t0

    def submit(self, req):
        self.waiting.append(req)

    def _admit(self):
        """Admit waiting requests in arrival order while the HBM budget has room."""
        while self.waiting and self.waiting[0].arrival <= self.now():
            r = self.waiting[0]
            slot = self.pool.add(**r.layers, rows=self.rows)
            if slot is None:
                if not self.active:
                    raise MemoryError(f"request {r.id} ({r.cells} cells) does not fit the HBM budget alone")
                return
            self.waiting.popleft()
            self.pool.set_state(slot, np.zeros(r.cells))
            r.slot = slot
            r.admitted = self.now()
            self.active.append(r)

    def step(self):
        """One iteration: admit, advance every active request by one chunk, retire finished ones."""
        self._admit()
        if not self.active:
            return False
        slots, targets = [], []
        for r in self.active:
            t1 = min(r.duration, r.t + self.chunk)
            self.pool.set_rain(r.slot, r.rain_fn(0.5 * (r.t + t1)))
            slots.append(r.slot)
            targets.append(t1)
        clocks = self.pool.advance(slots, targets, self.K)
        self.iterations += 1
        still = []
        for r, t1, c in zip(self.active, targets, clocks):
            self.cell_updates += c.cellUpdates - r.meta.get("_updates", 0)
            r.meta["_updates"] = c.cellUpdates
            r.t = t1
            if r.t >= r.duration:
                st = self.pool.read_state(r.slot, fields=self.keep)
                r.result = dict(steps=c.steps, mass=st["mass"], **{k: st[k] for k in self.keep})
                r.done = self.now()
                self.pool.remove(r.slot)
                r.slot = None
                self.finished.append(r)
            else:
                still.append(r)
        self.active = still
        return True

    def run(self, requests, idle_sleep=0.002):
        """Serve a list of requests (arrival times on the server clock) until all are done."""
        for r in sorted(requests, key=lambda q: q.arrival):
            self.submit(r)
        while self.waiting or self.active:
            if not self.step():
                nxt = self.waiting[0].arrival if self.waiting else None
                if nxt is not None:
                    time.sleep(max(0.0, min(idle_sleep * 50, nxt - self.now())))
        return self.finished

    def stats(self):
        lat = np.array([r.done - r.arrival for r in self.finished])
        span = max(r.done for r in self.finished) - min(r.arrival for r in self.finished)
        return dict(requests=len(self.finished), p50_s=float(np.percentile(lat, 50)),
                    p99_s=float(np.percentile(lat, 99)), mean_s=float(lat.mean()), span_s=float(span),
                    solves_per_hour=3600.0 * len(self.finished) / span,
                    cell_updates_per_s=self.cell_updates / span, iterations=self.iterations)

    def close(self):
        self.pool.destroy()


def solve_alone(req, chunk_s=60.0, K=64, rows=64, keep=("hmax",)):
    """The same request through the same chunk schedule in a pool of its own."""
    S = Server(0, chunk_s=chunk_s, K=K, rows=rows, clock=lambda: 0.0, keep=keep)
    try:
        r = Request(req.id, 0.0, req.layers, req.duration, req.rain_fn, dict(req.meta))
        S.run([r])
        return r.result
    finally:
        S.close()

"""Batched solves: many uneven urban instances advanced by one launch sequence, and the tile-shape
autotuner.

The reference runs solves one after another (src/batch.js:1-4 drives one solver per job). Here a
`Pool` holds every instance device-resident; one captured K-step graph advances all of them, each
with its own dt, clock and FP64 books, and the strips of every instance share each launch. The
tile shape is the number of rows a wavefront strip walks (the strip is 58 columns wide); results
do not depend on it, only the time does, so the autotuner picks it by measurement.

    specs = instances(ensemble(256))                 # urban layers per instance
    out = run_batched(specs, t_end=300.0)            # -> dict(wall_s, states, rows, ...)
    ref = run_sequential(specs, t_end=300.0)         # one instance at a time
"""
import time

import numpy as np

from . import scenarios as sc
from .solver import Pool

ROW_CANDIDATES = (16, 32, 64, 128)


def instance(spec, base_rain_mmh=40.0):
    """Layers of one ensemble member (scenarios.ensemble entry): the urban generator at its side
    and seed, Manning n scaled, all edges free outfall, and its rain rate (m/s)."""
    u = sc.urban(spec["side"], seed=spec["seed"])
    return dict(nx=u["nx"], ny=u["ny"], dx=u["dx"], z=u["z"], mask=u["mask"],
                manning=u["manning"] * spec["n_scale"],
                bc={k: "outflow" for k in ("west", "east", "south", "north")},
                rain=base_rain_mmh * spec["rain_mult"] * sc.MMHR, seed=spec["seed"])


def instances(ensemble_specs, base_rain_mmh=40.0):
    return [instance(s, base_rain_mmh) for s in ensemble_specs]


def _layers(inst):
    return {k: inst[k] for k in ("nx", "ny", "dx", "z", "mask", "manning", "bc")}


def add_all(pool, insts, rows=None):
    slots = []
    for inst in insts:
        s = pool.add(**_layers(inst), rows=rows or 0)
        if s is None:
            raise MemoryError("pool budget too small for the batch")
        pool.set_state(s, np.zeros(inst["nx"] * inst["ny"]))
        pool.set_rain(s, inst["rain"])
        slots.append(s)
    return slots


def tune_rows(pool, slots, targets, candidates=ROW_CANDIDATES, steps=4, repeats=2):
    """Autotuner: time `steps` steps of the listed instances for each candidate strip height and
    keep the fastest. The measured steps are real steps of the solve (nothing is thrown away),
    and the choice does not change any result. Returns (best rows, {rows: ms per step})."""
    timing = {}
    for r in candidates:
        for s in slots:
            pool.set_rows(s, r)
        pool.step(slots, targets, 1)                 # rebuild the group and capture its graph
        best = float("inf")
        for _ in range(repeats):
            pool.clock(slots[0])                     # synchronise
            t0 = time.perf_counter()
            pool.step(slots, targets, steps)
            best = min(best, (time.perf_counter() - t0) / steps)
        timing[r] = best * 1e3
    rows = min(timing, key=timing.get)
    for s in slots:
        pool.set_rows(s, rows)
    return rows, timing


def run_batched(insts, t_end, K=64, budget=0, tune=True, read=True):
    """All instances in one pool, advanced together to t_end. Returns wall time (s, tuning
    included), the tuned rows, per-instance states (if read) and steps."""
    P = Pool(budget)
    try:
        slots = add_all(P, insts)
        targets = [t_end] * len(slots)
        t0 = time.perf_counter()
        rows, timing = tune_rows(P, slots, targets) if tune else (64, {})
        P.advance(slots, targets, K)
        P.clock(slots[0])
        wall = time.perf_counter() - t0
        clocks = [P.clock(s) for s in slots]
        states = [P.read_state(s) for s in slots] if read else None
        return dict(wall_s=wall, rows=rows, tune_ms=timing, states=states,
                    steps=[c.steps for c in clocks], cell_updates=sum(c.cellUpdates for c in clocks))
    finally:
        P.destroy()


def _size_class(inst):
    return int(np.ceil(np.log2(max(inst["nx"], inst["ny"]))))


def run_sequential(insts, t_end, K=64, tune=True, read=False):
    """The plain form: one instance at a time (its own pool), each with the tile shape the
    autotuner picked for its size class (tuned once per class, its tuning time included)."""
    tuned = {}
    wall = 0.0
    states, steps, updates = [], [], 0
    for inst in insts:
        P = Pool(0)
        try:
            (s,) = add_all(P, [inst])
            t0 = time.perf_counter()
            cls = _size_class(inst)
            if tune and cls not in tuned:
                tuned[cls], _ = tune_rows(P, [s], [t_end])
            P.set_rows(s, tuned.get(cls, 64))
            P.advance([s], [t_end], K)
            c = P.clock(s)
            wall += time.perf_counter() - t0
            steps.append(c.steps)
            updates += c.cellUpdates
            if read:
                states.append(P.read_state(s))
        finally:
            P.destroy()
    return dict(wall_s=wall, rows=tuned, states=states if read else None, steps=steps, cell_updates=updates)

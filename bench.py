"""Benchmark of the shallow-water step on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload urban8192|urban2048|dambreak256] [--gpus 1]

Default workload: the 8192 x 8192 synthetic urban domain at 1 m (0.5 % plane + fBm, ~30 % building
walls, paved/pervious n, rain 50 mm/h, a triangular point inflow 0 -> 40 m^3/s at the west edge,
free-outfall edges) with local time stepping on 4 levels, spun up from dry for 600 s of simulated
time in 10 s frames, then W untimed and K timed steps. With local time stepping a step is one
cycle (2^4 base steps, 31 level records); with --stepping global it is one SSP-RK2 step. The
state stays resident in HBM throughout.

Prints ONE JSON line: cell-updates/s (open cells advanced per second), the roofline of the
dominant kernel (second RK stage + sources) measured with HIP events over the timed region, and
the FP64 CPU oracle timed on a bounded window of the same state as the CPU baseline. Two more
legs follow (skip with --no-batch / --no-serve): `batched`, 256 uneven urban instances (sides
128-1024) solved for 300 s from dry in one pool vs one at a time (solves/hour), with the smallest
checked against the oracle; `served`, a Poisson stream (2 requests/s, 32 requests, sides
256-4096, 30-minute storms) under continuous batching (p50/p99 latency, solves/hour, against one at a time).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, vendor figure
B_STAGE1 = 41                  # algorithmic bytes per grid cell, first stage: U^n 16 + z,n,mask 9 | U* 16
B_STAGE2 = 65                  # second stage: U^n 16 + U* 16 + static 9 + hmax 4 | U^{n+1} 16 + hmax 4
B_STEP = B_STAGE1 + B_STAGE2   # 106 B per cell-update (FP32 state with the FP32 depth residual)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_of(workload, stepping):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary of this workload (calibrated against the copy kernel), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        e = t[f"{workload}/{stepping}"]
        return e["bytes_per_launch"], e
    except (OSError, KeyError, ValueError):
        return None, None


def build_workload(name, seed=2026):
    from floodsim_amd import scenarios as sc
    if name.startswith("urban"):
        return dict(sc.urban_inflow(int(name[5:]), seed=seed), hydro=True, lts=4)
    if name == "dambreak256":
        d = sc.dam_break(256)
        return dict(nx=256, ny=256, dx=1.0, z=d["z"], mask=None, manning=0.03, bc={}, inlets=None, rain=0.0,
                    hydro=False, depth=d["depth"], cfl=0.5, lts=0)
    raise SystemExit(f"unknown workload {name}")


def cpu_threads():
    """Host threads for the all-cores baseline: every core the process may use, at most 16 (one GPU's share of
    the host on the GPU machines)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline(w, state, lts=0, window=1024, steps=None, budget_s=20.0, threads=1):
    """The FP64 oracle (same stepping) on a window of the same spun-up state, on `threads` OpenMP threads."""
    import oracle
    oracle.set_threads(threads)
    n = w["nx"]
    m = min(window, n)
    j0 = i0 = (n - m) // 2
    sl = (slice(j0, j0 + m), slice(i0, i0 + m))
    cut = lambda a: np.ascontiguousarray(np.asarray(a).reshape(n, n)[sl]).reshape(-1)
    man = w["manning"] if np.isscalar(w["manning"]) else cut(w["manning"])
    c = oracle.CPUSolver(nx=m, ny=m, dx=w["dx"], z=cut(w["z"]), manning=man,
                         mask=None if w["mask"] is None else cut(w["mask"]), bc=w["bc"], hdry=1e-5, lts=lts)
    c.setState(cut(state["h"]), cut(state["hu"]), cut(state["hv"]), t=0.0)
    c.rainRate = w["rain"]
    open_cells = m * m if w["mask"] is None else int(cut(w["mask"]).sum())
    # one step to size the sample, then as many as fit the budget
    t0 = time.perf_counter()
    c.step()
    one = time.perf_counter() - t0
    k = steps or max(1, min(200, int(budget_s / max(one, 1e-6))))
    u0 = c.cellSteps
    t0 = time.perf_counter()
    for _ in range(k):
        c.step()
    el = time.perf_counter() - t0
    upd = (c.cellSteps - u0) if lts else open_cells * k
    kind = "LTS cycles (%d levels)" % lts if lts else "global SSP-RK2 steps"
    oracle.set_threads(1)
    return dict(value=upd / el, unit="cell-updates/s", cores=threads, kind="port", steps=k,
                sample=f"FP64 oracle (C restatement of solver-cpu.js), {threads} thread{'s' if threads > 1 else ''}, "
                       f"{m}x{m} centre window of the spun-up state, {k} {kind} in {el:.1f} s")


def bench_batched(count=256, t_end=300.0, seed=7):
    """Config 4: `count` uneven urban instances (sides log-uniform 128-1024, own seeds, rain 40 mm/h
    x LogNormal(0, 0.3), Manning n x U(0.8, 1.2)) solved from dry for t_end simulated seconds, all
    in one pool (autotuned strip rows), then one at a time; the smallest instance is checked
    against the FP64 oracle."""
    import oracle
    from floodsim_amd import batch
    from floodsim_amd import scenarios as sc
    t0 = time.perf_counter()
    specs = sc.ensemble(count, seed=seed)
    insts = batch.instances(specs)
    cells = [i["nx"] * i["ny"] for i in insts]
    log(f"batched: {count} instances, {sum(cells)} cells, generated in {time.perf_counter() - t0:.1f} s")
    b = batch.run_batched(insts, t_end, read=False)
    log(f"batched: {b['wall_s']:.2f} s (rows {b['rows']})")
    q = batch.run_sequential(insts, t_end)
    log(f"one at a time: {q['wall_s']:.2f} s (rows per size class {q['rows']})")
    # parity sample against the FP64 oracle: the smallest and the largest instance and two drawn at random, each
    # re-solved in the batch's own pool form (the same bits as in the batch: tests/test_gpu_batch.py)
    rng = np.random.default_rng(seed)
    order = np.argsort(cells)
    sample = [int(order[0]), int(order[-1])] + [int(i) for i in rng.choice(order[1:-1], size=2, replace=False)]
    oracle.set_threads(cpu_threads())
    par = []
    for k in sample:
        inst = insts[k]
        P = batch.Pool(0)
        slots = batch.add_all(P, [inst])
        P.advance(slots, [t_end])
        g = P.read_state(slots[0])
        steps_g = P.clock(slots[0]).steps
        P.destroy()
        c = oracle.CPUSolver(nx=inst["nx"], ny=inst["ny"], dx=inst["dx"], z=inst["z"], mask=inst["mask"],
                             manning=inst["manning"], bc=inst["bc"], hdry=1e-5)
        c.setDepth(np.zeros(cells[k]))
        c.rainRate = inst["rain"]
        c.run(t_end)
        par.append({"instance": k, "side": inst["nx"], "max_dh_m": float(np.max(np.abs(g["h"] - c.field("h")))),
                    "mass_rel": g["mass"]["rel"], "steps_gpu": steps_g, "steps_oracle": c.steps})
        log(f"batched parity: instance {k} ({inst['nx']}^2) max|dh| {par[-1]['max_dh_m']:.2e} m")
    oracle.set_threads(1)
    return {"instances": count, "cells": int(sum(cells)), "sim_s": t_end, "seed": seed,
            "solves_per_hour": 3600.0 * count / b["wall_s"],
            "sequential_solves_per_hour": 3600.0 * count / q["wall_s"],
            "speedup_vs_one_at_a_time": q["wall_s"] / b["wall_s"],
            "batched_wall_s": b["wall_s"], "sequential_wall_s": q["wall_s"],
            "cell_updates_per_s": b["cell_updates"] / b["wall_s"], "rows_autotuned": b["rows"],
            "tuning_ms_per_step": b["tune_ms"],
            "parity_sample": par, "max_dh_m": max(p["max_dh_m"] for p in par)}


def bench_served(rate=2.0, count=32, budget_gb=64, seed=11):
    """Config 5: a Poisson stream (`rate` per s) of `count` urban solves, sides drawn from {256, ..., 4096} with
    p {.35, .3, .2, .1, .05}, each under the peaked 30-minute storm (120 mm/h at 40 %, 25 mm) scaled by
    LogNormal(0, 0.25), served with continuous batching under an HBM budget: p50/p99 latency and throughput of
    the stream. Then the same requests all waiting at once (saturated) through the server, and one at a time,
    each alone in its own pool through the same chunk schedule (the plain GPU form): solves/hour of both, and
    every request's result alone against its result in the mix."""
    from floodsim_amd import scenarios as sc
    from floodsim_amd import serve
    arr = sc.poisson_arrivals(rate, count, seed=seed)
    mk = lambda a, i, t: serve.storm_request(i, t, a["side"], a["seed"], a["depth_scale"])
    reqs = [mk(a, i, a["t"]) for i, a in enumerate(arr)]
    S = serve.Server(int(budget_gb) << 30, chunk_s=60.0)
    try:
        fin = S.run(reqs)
        st = S.stats()
    finally:
        S.close()
    log(f"served: {count} requests at {rate}/s, p50 {st['p50_s']:.2f} s, p99 {st['p99_s']:.2f} s")
    sat_reqs = [mk(a, i, 0.0) for i, a in enumerate(arr)]
    S = serve.Server(int(budget_gb) << 30, chunk_s=60.0)
    try:
        t0 = time.perf_counter()
        S.run(sat_reqs)
        sat_wall = time.perf_counter() - t0
        sat = S.stats()
    finally:
        S.close()
    log(f"served, all waiting at once: {sat_wall:.2f} s")
    seq_wall, same = 0.0, True
    for r in fin:
        t0 = time.perf_counter()
        alone = serve.solve_alone(r)
        seq_wall += time.perf_counter() - t0
        same = same and alone["steps"] == r.result["steps"] and bool(np.array_equal(alone["hmax"], r.result["hmax"]))
    log(f"one at a time: {seq_wall:.2f} s; alone = mixed for every request: {same}")
    st.update({"rate_per_s": rate, "budget_GiB": budget_gb, "seed": seed,
               "sides": {str(k): int(sum(a["side"] == k for a in arr)) for k in sorted({a["side"] for a in arr})},
               "storm": "30 min, peak 120 mm/h at 40 %, 25 mm, rates x LogNormal(0, 0.25) per request",
               "chunk_s": 60.0, "alone_equals_mixed": same, "requests_checked_alone": len(fin),
               "max_mass_rel": max(abs(r.result["mass"]["rel"]) for r in fin),
               "saturated": {"wall_s": sat_wall, "solves_per_hour": 3600.0 * count / sat_wall,
                             "cell_updates_per_s": sat["cell_updates_per_s"], "p50_s": sat["p50_s"],
                             "p99_s": sat["p99_s"]},
               "sequential_wall_s": seq_wall, "sequential_solves_per_hour": 3600.0 * count / seq_wall,
               "speedup_vs_one_at_a_time": seq_wall / sat_wall})
    return st


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (LTS: cycles); default 40 cycles / 300 steps")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps first; default 4 cycles / 20 steps")
    ap.add_argument("--workload", default="urban8192")
    ap.add_argument("--spinup", type=float, default=600.0)
    ap.add_argument("--stepping", choices=("lts", "global"), default=None)
    ap.add_argument("--rows", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--hazard", action="store_true", help="track the hazard products every step (as the reference does)")
    ap.add_argument("--no-timing", action="store_true",
                    help="no HIP events around the timed launches (for counter-collection runs)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched (256 instances) measurement")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-stream measurement")
    ap.add_argument("--serve-rate", type=float, default=2.0)
    ap.add_argument("--serve-count", type=int, default=32)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported (the path shards by "
              f"domain strips or by instance; the library has no multi-GPU path)", file=sys.stderr)
        sys.exit(2)

    import floodsim_amd as fsa
    from floodsim_amd import scenarios as sc
    from floodsim_amd.solver import copy_probe

    t_setup = time.perf_counter()
    w = build_workload(args.workload)
    log(f"workload {args.workload}: generated in {time.perf_counter() - t_setup:.1f} s")
    stepping = args.stepping or ("lts" if w.get("lts") else "global")
    lts = (w.get("lts") or 4) if stepping == "lts" else 0
    if args.steps is None:
        args.steps = 40 if lts else 300
    if args.warmup is None:
        args.warmup = 4 if lts else 20
    g = fsa.GPUSolver(nx=w["nx"], ny=w["ny"], dx=w["dx"], z=w["z"], manning=w["manning"], mask=w["mask"],
                      bc=w["bc"], inlets=w["inlets"], cfl=w.get("cfl"), rows=args.rows, lts=lts, hazard=args.hazard)
    if "depth" in w:
        g.setDepth(w["depth"])
    else:
        g.setDepth(lambda X, Y, Z: 0 * X)
    g.rainRate = w["rain"]
    # spin-up in 10 s frames (the reference app's quanta, app/map.js:2281); the hydrograph is held
    # per frame at its mid-frame value
    t = 0.0
    seg = 10.0
    while t < args.spinup - 1e-9:
        if w["hydro"]:
            g.setInletRates([-float(sc.triangle_hydrograph(t + seg / 2))])
        t = min(args.spinup, t + seg)
        g.advance(t, 8 if lts else 64)
    log(f"spun up to t={g.t:.1f} s in {g.steps} steps ({time.perf_counter() - t_setup:.1f} s total)")
    big = 1e9
    from floodsim_amd.solver import lib
    import ctypes
    L = lib()
    pool = L.fs_solver_pool(g._h)
    sl = (ctypes.c_int32 * 1)(0)
    tg = (ctypes.c_double * 1)(big)

    def run_steps(k):
        rc = L.fs_pool_step(pool, sl, tg, 1, int(k))
        if rc != 0:
            raise RuntimeError(L.fs_last_error(g._h).decode())

    c0 = g.clock()
    run_steps(args.warmup)
    if not args.no_timing:
        g.set_timing(True)
    c1 = g.clock()                                   # synchronises the stream
    t0 = time.perf_counter()
    run_steps(args.steps)
    c2 = g.clock()                                   # synchronises again
    el = time.perf_counter() - t0
    if args.no_timing:
        ms1 = ms2 = dom_ms = dom_cells = -1.0
        nt = dom_n = 0
    else:
        ms1, ms2, nt = g.kernel_ms()
        dom_ms, dom_cells, dom_n = g.dominant_kernel()
        g.set_timing(False)
    steps = c2.steps - c1.steps
    updates = c2.cellUpdates - c1.cellUpdates
    cells = w["nx"] * w["ny"]
    value = updates / el
    ms_step = 1e3 * el / max(steps, 1)
    # dominant kernel: the second RK stage (+ sources); algorithmic bytes = 65 B per open cell it advances
    achieved = dom_cells * B_STAGE2 / (dom_ms * 1e-3) / 1e9 if dom_ms > 0 else None
    s = g.readState()
    try:
        cp_gbs, _ = copy_probe(2 << 30, 10)
    except Exception as e:  # noqa: BLE001
        log("copy probe failed:", e)
        cp_gbs = None
    # (the committed summaries cover the plain kernel variant: none for the hazard-products variant)
    traffic, tr = traffic_of(args.workload, stepping + ("+hazard" if args.hazard else ""))
    cpu = None
    if not args.no_cpu:
        # one core (the reference's solver-cpu.js is single-threaded), then the same sample on every core
        one = cpu_baseline(w, s, lts=lts)
        nt = cpu_threads()
        allc = cpu_baseline(w, s, lts=lts, steps=one["steps"], threads=nt) if nt > 1 else one
        cpu = dict(allc, host_cpus=os.cpu_count(), single_core=one,
                   speedup_all_cores=allc["value"] / one["value"])
        log(f"cpu baseline: {one['value']:.3e} cell-updates/s on 1 core, {allc['value']:.3e} on {nt}")
    g.destroy()
    batched = None if args.no_batch else bench_batched()
    served = None if args.no_serve else bench_served(args.serve_rate, args.serve_count)
    out = {
        "metric": "cell-updates/s + % HBM roofline; batched solves/hr, p99 latency @ \u22641 mm",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 books)",
        "data": "synthetic",
        "config": {"workload": args.workload, "nx": w["nx"], "ny": w["ny"], "dx_m": w["dx"],
                   "open_cells": int(cells if w["mask"] is None else int(np.asarray(w["mask"]).sum())),
                   "stepping": ("LTS, %d levels (a step is one cycle)" % lts) if lts else "global SSP-RK2",
                   "spinup_s": args.spinup, "t_start": c1.t, "t_end": c2.t,
                   "layers_checksum": w.get("checksum"), "rows_per_strip": args.rows or 64,
                   "hazard_products": bool(args.hazard)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
                     "traffic_source": tr and tr.get("source"),
                     "kernel": "fs::sw_stage<1,%s%s> (second RK stage + sources)" % ("true" if lts else "false",
                                                                                ",hazard" if args.hazard else ""),
                     "bytes_per_cell_update": B_STAGE2, "launches": dom_n,
                     "cells_per_launch": dom_cells / max(dom_n, 1), "avg_ms": dom_ms / max(dom_n, 1),
                     "copy_probe_GBps": cp_gbs, "frac_of_copy": achieved / cp_gbs if achieved and cp_gbs else None,
                     "traffic_GBps": traffic / (dom_ms / max(dom_n, 1) * 1e-3) / 1e9 if traffic and dom_ms > 0 else None},
        "cpu_baseline": cpu,
        "batched": batched,
        "served": served,
        "extra": {"stage1_ms": ms1, "stage2_ms": ms2, "timed_launches": nt,
                  "step_GBps_106B": value * B_STEP / 1e9,
                  "step_frac_106B": value * B_STEP / 1e9 / HBM_PEAK_GBS,
                  "copy_probe_GBps": cp_gbs, "all_cell_updates_per_s": cells * steps / el,
                  "mass_rel": s["mass"]["rel"], "last_dt": c2.lastDt},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Batched and served solves: every instance gives the same bits alone and in the mix, a sample
matches the FP64 oracle, admission keeps to the HBM budget, and the autotuned tile shape changes
the time only."""
import numpy as np
import pytest

import oracle
from floodsim_amd import scenarios as sc

pytestmark = pytest.mark.gpu

TOL_H = 1e-3
TOL_MASS = 1e-6


def _specs():
    sides = [96, 130, 181, 72, 240, 110]
    rng = np.random.default_rng(3)
    return [dict(side=s, seed=int(rng.integers(1, 2 ** 31 - 1)), rain_mult=float(rng.lognormal(0, 0.3)),
                 n_scale=float(rng.uniform(0.8, 1.2))) for s in sides]


def test_batched_equals_alone_and_oracle():
    from floodsim_amd import batch
    insts = batch.instances(_specs())
    T = 120.0
    b = batch.run_batched(insts, T)
    s = batch.run_sequential(insts, T, read=True)
    print(f"batched {b['wall_s']:.3f} s (rows {b['rows']}, tuning {b['tune_ms']}), sequential {s['wall_s']:.3f} s")
    assert b["steps"] == s["steps"]
    for a, m in zip(s["states"], b["states"]):
        for k in ("h", "hu", "hv", "hmax"):
            np.testing.assert_array_equal(a[k], m[k])
        assert abs(m["mass"]["rel"]) <= TOL_MASS
    # a sample against the FP64 oracle
    for i in (0, 3):
        inst = insts[i]
        c = oracle.CPUSolver(nx=inst["nx"], ny=inst["ny"], dx=inst["dx"], z=inst["z"], mask=inst["mask"],
                             manning=inst["manning"], bc=inst["bc"], hdry=1e-5)
        c.setDepth(np.zeros(inst["nx"] * inst["ny"]))
        c.rainRate = inst["rain"]
        c.run(T)
        d = float(np.max(np.abs(b["states"][i]["h"] - c.field("h"))))
        print(f"instance {i} ({inst['nx']}^2): {b['steps'][i]} steps (oracle {c.steps}), max|dh| {d:.2e} m")
        assert d <= TOL_H


def test_rows_change_time_not_results():
    from floodsim_amd.solver import Pool
    from floodsim_amd import batch
    inst = batch.instances(_specs()[:1])[0]
    out = []
    for rows in (8, 64):
        P = Pool()
        (s,) = batch.add_all(P, [inst])
        P.set_rows(s, rows)
        P.advance([s], [60.0])
        P.set_rows(s, 200 - rows)        # change mid-solve as the autotuner does
        P.advance([s], [90.0])
        out.append(P.read_state(s))
        P.destroy()
    for k in ("h", "hu", "hv", "hmax"):
        np.testing.assert_array_equal(out[0][k], out[1][k])


def test_served_alone_vs_mix_and_budget():
    from floodsim_amd import serve
    from floodsim_amd.solver import Pool
    sides = [64, 120, 90, 150, 70]
    reqs = [serve.storm_request(i, arrival=float(i), side=s, seed=100 + i, depth_scale=1.0 + 0.1 * i,
                                duration=240.0) for i, s in enumerate(sides)]
    need = sorted(Pool.bytes_needed(r.layers["nx"], r.layers["ny"]) for r in reqs)
    budget = need[-1] + need[-2] + 4096          # room for two of the largest at once
    tick = {"t": 0.0}

    def clock():
        return tick["t"]

    S = serve.Server(budget, chunk_s=60.0, clock=clock)
    for r in reqs:
        S.submit(r)
    peak = 0
    most = 0
    while S.waiting or S.active:
        S.step()
        peak = max(peak, S.pool.bytes_used())
        most = max(most, len(S.active))
        tick["t"] += 1.0                        # one iteration = one second of server time
    st = S.stats()
    S.close()
    print(f"served {st['requests']} requests, p50 {st['p50_s']:.1f}, p99 {st['p99_s']:.1f} (virtual s), "
          f"peak bytes {peak} of budget {budget}")
    assert st["requests"] == len(reqs)
    assert peak <= budget
    # continuous batching: requests ran together, and the budget held some back
    assert most >= 2
    assert any(r.admitted > r.arrival for r in reqs)
    for r in reqs:
        alone = serve.solve_alone(r)
        assert alone["steps"] == r.result["steps"]
        np.testing.assert_array_equal(alone["hmax"], r.result["hmax"])
        assert abs(r.result["mass"]["rel"]) <= TOL_MASS


def test_ensemble_sample_vs_oracle():
    """Config 4's own ensemble (256 instances, sides round(128 * 8^u), seed 7): 48 of its members, the largest
    (side 1014, the top of the 128-1024 range this draw reaches) and a random draw of the rest, batched together for 120 s; the largest and three random
    members against the FP64 oracle, every member's books within 1e-6."""
    from floodsim_amd import batch
    specs = sc.ensemble(256, seed=7)
    sides = np.array([s["side"] for s in specs])
    big = int(np.argmax(sides))
    rng = np.random.default_rng(17)
    rest = [int(i) for i in rng.choice(np.delete(np.arange(256), big), size=47, replace=False)]
    pick = [big] + rest
    insts = batch.instances([specs[i] for i in pick])
    T = 120.0
    print(f"\nensemble sample: members {pick[:8]}..., largest side {sides[big]}, cells {sum(i['nx'] * i['ny'] for i in insts)}",
          flush=True)
    assert sides[big] >= 1000
    b = batch.run_batched(insts, T)
    for st in b["states"]:
        assert abs(st["mass"]["rel"]) <= TOL_MASS and st["mass"]["clipped"] == 0
    for j in [0] + [int(q) for q in rng.choice(np.arange(1, len(pick)), size=3, replace=False)]:
        inst = insts[j]
        c = oracle.CPUSolver(nx=inst["nx"], ny=inst["ny"], dx=inst["dx"], z=inst["z"], mask=inst["mask"],
                             manning=inst["manning"], bc=inst["bc"], hdry=1e-5)
        c.setDepth(np.zeros(inst["nx"] * inst["ny"]))
        c.rainRate = inst["rain"]
        c.run(T)
        d = float(np.max(np.abs(b["states"][j]["h"] - c.field("h"))))
        print(f"member {pick[j]} ({inst['nx']}^2): {b['steps'][j]} steps (oracle {c.steps}), max|dh| {d:.2e} m",
              flush=True)
        assert d <= TOL_H
        assert abs(b["steps"][j] - c.steps) <= 1

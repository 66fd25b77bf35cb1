"""Batched, served and checkpointed solves through Node: the N-API pool entries with the JavaScript Pool
and Server (floodsim_amd/js/pool-hip.js), and the checkpoint adapter (floodsim_amd/js/checkpoint-hip.js).

CPU: the modules load under Node, parse as ES2019, and export what their headers list.
GPU: the same instances through the JS Pool and the Python Pool give the same bits; the same served stream
through the JS Server and the Python Server admits and retires each request at the same iteration with the
same bits; a run checkpointed with gpuSnapshot / solverState and resumed in a new solver with gpuRestore /
applySolverState continues bit for bit (the reference app's use of src/checkpoint.js, app/map.js:2616-2660),
and gpuRestore size-checks every part before writing (checkpoint.js:106-112).
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")

pytestmark = pytest.mark.skipif(NODE is None, reason="node is not installed")


def _node_eval(script):
    return subprocess.run([NODE, "-e", script], cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                          timeout=120)


def _run_js(driver, case, tmp_path, name):
    p = tmp_path / name
    p.write_text(json.dumps(case))
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", driver), str(p)], cwd=ROOT, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


def test_pool_and_checkpoint_modules_load():
    r = _node_eval("const P=require('./floodsim_amd/js/pool-hip.js'); const C=require('./floodsim_amd/js/checkpoint-hip.js');"
                   "console.log(JSON.stringify([Object.keys(P).sort(), Object.keys(C).sort()]))")
    assert r.returncode == 0, r.stderr
    pool, ck = json.loads(r.stdout)
    assert pool == sorted(["Pool", "Server", "Request", "addAll", "tuneRows", "runBatched", "solveAlone", "ROW_CANDIDATES"])
    assert ck == sorted(["gpuSnapshot", "gpuRestore", "solverState", "applySolverState"])
    for f in ("floodsim_amd/js/pool-hip.js", "floodsim_amd/js/checkpoint-hip.js", "floodsim_amd/js/params.js",
              "tests/js/run_pool.js", "tests/js/run_checkpoint.js"):
        c = subprocess.run([NODE, "--check", os.path.join(ROOT, f)], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                           text=True)
        assert c.returncode == 0, f + ": " + c.stderr


def test_checkpoint_state_round_trip_without_device():
    """solverState / applySolverState on a plain object with the solver's JS fields (no device needed)."""
    r = _node_eval("const C=require('./floodsim_amd/js/checkpoint-hip.js');"
                   "const a={t:12.5, steps:40, lastDt:0.3, dtAvg:0.31, stage:99.5, vol:{rain:2, initial:1, exchanged:-3},"
                   " get rainRate(){return this._rain;}, _rain:1e-5};"
                   "const st=JSON.parse(JSON.stringify(C.solverState(a)));"
                   "const b={vol:{}, _rain:0, get rainRate(){return this._rain;}}; C.applySolverState(b, st);"
                   "console.log(JSON.stringify([b.t, b.steps, b.lastDt, b.dtAvg, b.stage, b.vol, b.rainRate, b.vol !== a.vol]))")
    assert r.returncode == 0, r.stderr
    assert json.loads(r.stdout) == [12.5, 40, 0.3, 0.31, 99.5, {"rain": 2, "initial": 1, "exchanged": -3}, 1e-5, True]


def _layers_json(inst):
    return dict(nx=int(inst["nx"]), ny=int(inst["ny"]), dx=float(inst["dx"]), z=np.asarray(inst["z"], float).tolist(),
                mask=np.asarray(inst["mask"], np.uint8).reshape(-1).tolist(),
                manning=np.asarray(inst["manning"], float).reshape(-1).tolist(), bc=inst["bc"])


def _layers_py(inst):
    return {k: inst[k] for k in ("nx", "ny", "dx", "z", "mask", "manning", "bc")}


@pytest.mark.gpu
def test_js_pool_equals_python_pool(tmp_path):
    """Four uneven urban instances (sides 96-181, their own rain and roughness) in one pool, each with its own
    strip height, changed mid-solve as the autotuner does: the JS Pool and the Python Pool give the same steps,
    fields and books."""
    from floodsim_amd import batch
    from floodsim_amd.solver import Pool
    rng = np.random.default_rng(11)
    specs = [dict(side=s, seed=int(rng.integers(1, 2 ** 31 - 1)), rain_mult=float(rng.lognormal(0, 0.3)),
                  n_scale=float(rng.uniform(0.8, 1.2))) for s in (96, 130, 72, 181)]
    insts = batch.instances(specs)
    rows, rows2 = [16, 64, 32, 128], [128, 16, 64, 32]
    t1, T = 40.0, 90.0
    case = dict(batch=dict(t1=t1, T=T, instances=[dict(_layers_json(i), rain=float(i["rain"]), rows=r, rows2=r2)
                                                   for i, r, r2 in zip(insts, rows, rows2)]))
    js = _run_js("run_pool.js", case, tmp_path, "batch.json")["batch"]
    P = Pool()
    slots = []
    for inst, r in zip(insts, rows):
        s = P.add(**_layers_py(inst), rows=r)
        P.set_state(s, np.zeros(inst["nx"] * inst["ny"]))
        P.set_rain(s, inst["rain"])
        slots.append(s)
    P.advance(slots, [t1] * len(slots))
    for s, r2 in zip(slots, rows2):
        P.set_rows(s, r2)
    clocks = P.advance(slots, [T] * len(slots))
    py = [P.read_state(s) for s in slots]
    used = P.bytes_used()
    P.destroy()
    print(f"JS pool: steps {js['steps']}, Python {[c.steps for c in clocks]}")
    assert js["steps"] == [c.steps for c in clocks]
    assert js["t"] == [c.t for c in clocks] == [T] * len(slots)
    assert js["bytesUsed"] == used
    for a, b in zip(js["states"], py):
        for k in ("h", "hu", "hv", "hmax"):
            assert np.array_equal(np.array(a[k]), b[k]), k
        for k in ("volume", "rain", "outflow", "expected", "rel"):
            assert a["mass"][k] == b["mass"][k], k
        assert abs(a["mass"]["rel"]) <= 1e-6


@pytest.mark.gpu
def test_js_server_equals_python_server(tmp_path):
    """Five storm requests (sides 64-150, arriving 1 s apart on a virtual clock that moves 1 s per iteration)
    under a budget that holds two of the largest at once: the JS Server and the Python Server admit and retire
    each request at the same iteration, and every result has the same steps and peak depths; the budget holds
    and requests wait for room."""
    from floodsim_amd import serve
    from floodsim_amd.solver import Pool
    sides = [64, 120, 90, 150, 70]
    chunk = 60.0
    reqs = [serve.storm_request(i, arrival=float(i), side=s, seed=100 + i, depth_scale=1.0 + 0.1 * i, duration=240.0)
            for i, s in enumerate(sides)]
    need = sorted(Pool.bytes_needed(r.layers["nx"], r.layers["ny"]) for r in reqs)
    budget = need[-1] + need[-2] + 4096

    def table(r):
        out, t = [], 0.0
        while t < r.duration:
            t1 = min(r.duration, t + chunk)
            mid = 0.5 * (t + t1)
            out.append([mid, float(r.rain_fn(mid))])
            t = t1
        return out

    case = dict(serve=dict(budget=budget, chunk=chunk, requests=[
        dict(id=r.id, arrival=r.arrival, duration=r.duration, layers=_layers_json(r.layers), rain=table(r)) for r in reqs]))
    js = _run_js("run_pool.js", case, tmp_path, "serve.json")["serve"]
    tick = {"t": 0.0}
    S = serve.Server(budget, chunk_s=chunk, clock=lambda: tick["t"])
    for r in reqs:
        S.submit(r)
    while S.waiting or S.active:
        S.step()
        tick["t"] += 1.0
    iters = S.iterations
    S.close()
    print(f"JS server: {js['iterations']} iterations (Python {iters}), peak bytes {js['peak']} of {budget}, "
          f"most active {js['most']}")
    assert js["iterations"] == iters
    assert js["peak"] <= budget and js["most"] >= 2
    assert any(q["admitted"] > r.arrival for q, r in zip(js["requests"], reqs))
    for q, r in zip(js["requests"], reqs):
        assert q["id"] == r.id and q["admitted"] == r.admitted and q["done"] == r.done
        assert q["steps"] == r.result["steps"]
        assert np.array_equal(np.array(q["hmax"]), r.result["hmax"])
        assert q["mass"]["volume"] == r.result["mass"]["volume"]
        assert abs(q["mass"]["rel"]) <= 1e-6


def _rich_case(lts):
    """Every part a snapshot can hold: flood barriers, Green-Ampt soil and GI stores, stage cells, and a rain
    field given after creation (the case of test_gpu_snapshot's rich solver)."""
    from oracle import refcases as rc
    sc = rc.cloudburst(80)
    nx, ny = sc["nx"], sc["ny"]
    n = nx * ny
    z = np.asarray(sc["z"], dtype=np.float64).reshape(-1)
    r = np.random.default_rng(9)
    ga = dict(psiD=np.full(n, 0.05), fmax=np.full(n, 0.3), perv=np.where(r.random(n) < 0.5, 0.6, 0.0))
    gi = dict(rate=np.where(r.random(n) < 0.2, 5e-5, 0.0), cap=np.full(n, 0.05), drain=np.full(n, 1e-6))
    crest = [dict(i=40, j=j, axis=0, z=float(max(z[j * nx + 39], z[j * nx + 40])) + 0.02, C=1.6) for j in range(10, 30)]
    man = np.broadcast_to(np.asarray(sc["manning"], dtype=np.float64), (n,)).copy() if np.isscalar(sc["manning"]) \
        else np.asarray(sc["manning"], dtype=np.float64).reshape(-1)
    infil = np.broadcast_to(np.asarray(sc["infil"], dtype=np.float64), (n,)).copy() if np.isscalar(sc["infil"]) \
        else np.asarray(sc["infil"], dtype=np.float64).reshape(-1)
    return dict(nx=nx, ny=ny, dx=float(sc["dx"]), z=z, mask=np.asarray(sc["mask"], np.uint8).reshape(-1), manning=man,
                bc=sc["bc"], lts=lts, infil=infil, stageCells=[j * nx for j in range(ny)], crest=crest, ga=ga, gi=gi,
                stage=float(z.min()) + 0.5, rainField=np.where(np.arange(n) % 4 == 0, 0.0, 1.5), rain=120 * rc.MMHR,
                frames=[30.0, 90.0], tEnd=200.0)


def _json(c):
    out = {}
    for k, v in c.items():
        if isinstance(v, np.ndarray):
            out[k] = v.tolist()
        elif isinstance(v, dict):
            out[k] = {q: (w.tolist() if isinstance(w, np.ndarray) else w) for q, w in v.items()}
        else:
            out[k] = v
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("lts", [0, 3])
def test_js_checkpoint_resume_identity(tmp_path, lts):
    import floodsim_amd as fsa
    c = _rich_case(lts)
    js = _run_js("run_checkpoint.js", _json(c), tmp_path, "ck.json")
    assert js["names"] == ["record", "planes", "barriers", "soil", "rainField"]
    assert js["sizes"] == js["bufSizes"] and min(js["sizes"]) > 0
    assert js["tB"] == js["tSnap"] == c["frames"][-1] and js["stepsB"] == js["stepsSnap"]
    assert js["rainB"] == c["rain"]
    # resumed = uninterrupted, bit for bit, books and barrier books included
    assert js["stepsBEnd"] == js["stepsA"]
    for k in ("h", "hu", "hv", "hmax"):
        assert js["a"][k] == js["b"][k], k
    for k in ("volume", "rain", "infil", "outflow", "giStored", "expected"):
        assert js["a"]["mass"][k] == js["b"]["mass"][k], k
    assert js["sameBars"] and js["volA"] == js["volB"]
    assert js["a"]["mass"]["giStored"] > 0 and js["a"]["mass"]["infil"] > 0
    # the size check (checkpoint.js:108-109) and a part left out
    assert js["errGrid"].startswith("checkpoint: solver buffer planes is ") and "(a different grid or options)" in js["errGrid"]
    assert js["errMissing"] == "checkpoint: solver buffer soil is not in the checkpoint"
    # and the uninterrupted JS run is the Python binding's
    s = fsa.GPUSolver(nx=c["nx"], ny=c["ny"], dx=c["dx"], z=c["z"], mask=c["mask"], manning=c["manning"], bc=c["bc"],
                      lts=lts, infil=c["infil"], stageCells=c["stageCells"], crest=c["crest"], ga=c["ga"], gi=c["gi"],
                      hazard=False)
    s.setStage(c["stage"])
    s.setRainField(c["rainField"])
    s.setDepth(lambda X, Y, Z: 0 * X)
    s.rainRate = c["rain"]
    for t in c["frames"] + [c["tEnd"]]:
        s.advance(t)
    py = s.readState()
    steps = s.steps
    s.destroy()
    assert steps == js["stepsA"]
    for k in ("h", "hu", "hv", "hmax"):
        assert np.array_equal(np.array(js["a"][k]), py[k]), k

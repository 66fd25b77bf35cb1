"""The Node boundary: N-API addon + the JavaScript GPUSolver (floodsim_amd/js/solver-hip.js).

CPU: the addon loads under Node and exports the functions the JS class uses; the class rejects
options of the reference outside the shallow-water step before touching the device.
GPU: one case driven through the JS class gives the same bits as the same case through the Python
binding (both call the one C ABI), and matches the FP64 oracle to 1 mm.
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "floodsim_amd", "js", "floodsim_napi.node")
NODE = shutil.which("node")

pytestmark = pytest.mark.skipif(NODE is None, reason="node is not installed")

ADDON_FUNCS = ["advance", "clock", "create", "debugNoSteps", "destroy", "frame", "levels", "readHazard", "readInlets", "readState",
               "readBarriers", "readInfiltration", "restore", "setCrestLevels", "setCrests", "setGa", "setGi", "setInletRates", "setRain",
               "setRainField", "setStage",
               "setStageCells", "setState", "snapshot", "snapshotLayout", "tileStats", "version",
               # the pool entries (fs_pool_*), for floodsim_amd/js/pool-hip.js
               "poolCreate", "poolDestroy", "poolBytesNeeded", "poolBytesUsed", "poolAdd", "poolRemove", "poolSetRows",
               "poolSetState", "poolSetRain", "poolSetInletRates", "poolSetStageCells", "poolSetStage", "poolSetCrests",
               "poolSetCrestLevels", "poolReadBarriers", "poolSetGa", "poolSetGi", "poolClock", "poolSetTiming",
               "poolKernelMs", "poolAdvance", "poolStep", "poolReadState"]


def node(script, *args, timeout=120):
    r = subprocess.run([NODE, "-e", script, *args], cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=timeout)
    return r


def test_addon_exports():
    assert os.path.exists(ADDON), "floodsim_napi.node is not built: run __graft_entry__.build()"
    r = node("const m=require('./floodsim_amd/js/floodsim_napi.node'); console.log(JSON.stringify(Object.keys(m)))")
    assert r.returncode == 0, r.stderr
    assert sorted(json.loads(r.stdout)) == sorted(ADDON_FUNCS)


OUT_OF_SCOPE = ["subgrid", "sink", "sediment", "debris", "scour", "banks", "underdrains", "lines"]


def test_js_class_loads_and_rejects_out_of_scope_options():
    r = node("const {GPUSolver, version}=require('./floodsim_amd/js/solver-hip.js'); console.log(version());"
             f"for (const k of {json.dumps(OUT_OF_SCOPE)}) {{ let msg='none';"
             " try { new GPUSolver(null, {nx:4, ny:4, dx:1, z:new Float64Array(16), [k]: [{}]}); }"
             " catch (e) { msg=e.message; } console.log(msg); }")
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert lines[0].startswith("floodsim_hip")
    assert len(lines) == 1 + len(OUT_OF_SCOPE)
    for k, msg in zip(OUT_OF_SCOPE, lines[1:]):
        assert k in msg and "not supported" in msg, (k, msg)


def test_js_sources_are_es2019():
    """Node 12 parses the class (no ??, ?., class fields): a syntax check of each JS file."""
    for f in ("floodsim_amd/js/solver-hip.js", "tests/js/run_case.js"):
        r = subprocess.run([NODE, "--check", os.path.join(ROOT, f)], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                           text=True)
        assert r.returncode == 0, f + ": " + r.stderr


def _case(nx=48, ny=40, seed=5):
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5)
    z = np.tile(0.01 * (nx - x), (ny, 1)) + 0.05 * rng.random((ny, nx))
    mask = np.ones((ny, nx), np.uint8)
    mask[10:16, 20:26] = 0
    depth = np.where(np.arange(nx)[None, :] < nx // 3, 0.6, 0.0) * np.ones((ny, 1))
    return dict(nx=nx, ny=ny, dx=1.0, z=z.reshape(-1).tolist(), manning=0.03, mask=mask.reshape(-1).tolist(),
                bc={"west": "wall", "east": "outflow", "south": "wall", "north": "wall"},
                depth=depth.reshape(-1).tolist(), rain=60 / 3.6e6, inlets=[20 * nx + 5], rates=[-0.2],
                frames=[2.0, 5.0, 8.0])


def _python_run(c):
    import floodsim_amd as fsa
    g = fsa.GPUSolver(nx=c["nx"], ny=c["ny"], dx=c["dx"], z=np.array(c["z"]), manning=c["manning"],
                      mask=np.array(c["mask"], np.uint8), bc=c["bc"], inlets=c["inlets"] or None,
                      stageCells=c.get("stageCells"), implicitFriction=c.get("implicitFriction", False),
                      stepFlux=c.get("stepFlux", False), hazard=c.get("hazard", True))
    g.setDepth(np.array(c["depth"]))
    g.rainRate = c["rain"]
    if c["rates"]:
        g.setInletRates(c["rates"])
    for f, t in enumerate(c["frames"]):
        if c.get("stageLevels"):
            g.setStage(c["stageLevels"][f])
        g.frame(t, 1 << 20, True)
    s = g.readState()
    steps = g.steps
    g.destroy()
    return s, steps


@pytest.mark.gpu
def test_js_matches_python_binding_and_oracle(tmp_path):
    import oracle
    c = _case()
    p = tmp_path / "case.json"
    p.write_text(json.dumps(c))
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "run_case.js"), str(p)], cwd=ROOT,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    js = json.loads(r.stdout)
    assert js["restoreSame"] and js["snapshotBytes"] > 0
    assert js["hazardCells"] > 0 and js["hazardForceMax"] > 0     # hazard products on by default, as the reference
    py, steps = _python_run(c)
    # the same C ABI under both bindings: identical bits
    assert js["steps"] == steps
    for k in ("h", "hu", "hv", "hmax"):
        assert np.array_equal(np.array(js[k]), py[k]), k
    assert js["mass"]["volume"] == py["mass"]["volume"]
    # against the FP64 oracle (1 mm), and mass closed on the GPU's own books
    o = oracle.CPUSolver(nx=c["nx"], ny=c["ny"], dx=1.0, z=np.array(c["z"]), manning=c["manning"],
                         mask=np.array(c["mask"], np.uint8), bc=c["bc"], inlets=c["inlets"], hdry=1e-5)
    o.setDepth(np.array(c["depth"]))
    o.rainRate = c["rain"]
    o.setInletRates(c["rates"])
    o.run(c["frames"][-1])
    d = float(np.max(np.abs(np.array(js["h"]) - o.field("h"))))
    print(f"JS path: {js['steps']} steps, max|dh| vs oracle {d:.2e} m, mass rel {js['mass']['rel']:.1e}")
    assert d <= 1e-3
    assert abs(js["mass"]["rel"]) <= 1e-6
    assert js["inletVol"] < 0            # water entered through the inlet (q < 0 is inflow)
    assert abs(js["volRain"] - js["mass"]["rain"]) <= 1e-6 * max(1.0, js["mass"]["rain"])


@pytest.mark.gpu
def test_js_stage_and_f26_case(tmp_path):
    """The water-level boundary and the F26 options through the JavaScript class: the sea edge of the
    reference's stage case (gpu-suite.js:148-164) rising 0.4 m, with implicit friction and the step flux;
    the same bits as the Python binding and within 1 mm of the FP64 oracle."""
    import oracle
    nx, ny, dx = 50, 20, 4.0
    I, J = np.meshgrid(np.arange(nx), np.arange(ny))
    z = (98 + 0.08 * I + 0.3 * np.sin(J)).reshape(-1)
    z0 = 99.5
    frames = [float(t) for t in range(30, 601, 30)]
    c = dict(nx=nx, ny=ny, dx=dx, z=z.tolist(), manning=0.02, mask=[1] * (nx * ny),
             bc={k: "outflow" for k in ("west", "east", "south", "north")}, depth=np.maximum(0, z0 - z).tolist(),
             rain=0.0, inlets=[], rates=[], frames=frames, stageCells=[j * nx for j in range(ny)],
             stageLevels=[z0 + 0.4 * min(1.0, t / 600) for t in frames], implicitFriction=True, stepFlux=True,
             hazard=False)
    p = tmp_path / "stage.json"
    p.write_text(json.dumps(c))
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "run_case.js"), str(p)], cwd=ROOT,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    js = json.loads(r.stdout)
    py, steps = _python_run(c)
    assert js["steps"] == steps
    for k in ("h", "hu", "hv", "hmax"):
        assert np.array_equal(np.array(js[k]), py[k]), k
    o = oracle.CPUSolver(nx=nx, ny=ny, dx=dx, z=z, manning=0.02, bc=c["bc"], stageCells=c["stageCells"],
                         implicitFriction=True, stepFlux=True, hdry=1e-5)
    o.setDepth(np.array(c["depth"]))
    for f, t in enumerate(frames):
        o.setStage(c["stageLevels"][f])
        o.run(t)
    d = float(np.max(np.abs(np.array(js["h"]) - o.field("h"))))
    print(f"JS stage case: {js['steps']} steps, max|dh| vs oracle {d:.2e} m, inflow {-js['mass']['outflow']:.2f} "
          f"m3 (oracle {-o.vol['outflow']:.2f}), mass rel {js['mass']['rel']:.1e}")
    assert d <= 1e-3 and abs(js["mass"]["rel"]) <= 1e-6 and -js["mass"]["outflow"] > 0


def test_ready_rejects_on_creation_error():
    """A creation error (here an invalid 1-column grid, or no device) does not throw from the constructor:
    `ready` rejects with 'GPU pipeline error: ...' (solver-gpu.js:287) and later calls fail with it."""
    r = node("const {GPUSolver}=require('./floodsim_amd/js/solver-hip.js');"
             "let s; try { s = new GPUSolver(null, {nx:1, ny:4, dx:1, z:new Float64Array(4)}); }"
             " catch (e) { console.log('threw ' + e.message); process.exit(0); }"
             "s.ready.then(() => console.log('resolved'), (e) => {"
             "  console.log(e.message);"
             "  return s.readState().then(() => console.log('readState resolved'), (e2) => console.log(e2.message)); });")
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert len(lines) == 2, lines
    assert lines[0].startswith("GPU pipeline error: ") and lines[1] == lines[0]
    assert "UnhandledPromiseRejection" not in r.stderr

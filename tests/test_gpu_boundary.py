"""Boundary behaviour of the solver object on the device, through the C ABI (Python) and the JS class:
- the no-progress error of advance (solver-gpu.js:601), forced from the host by the FS_DEBUG_NO_STEPS test
  switch (every launch group runs no steps, so no kernel is left waiting and the clock never moves);
- tileStats (solver-gpu.js:625-629) as tile counts: this build's skipping unit is the wavefront strip;
- readInlets().h (solver-gpu.js:564-576): the depth the last exchange left, cleared by the read as the
  reference clears its INX record;
- the 2^28-cell limit of the 16-byte per-cell planes (Green-Ampt soil, GI stores, barrier crests) on a
  global-stepping solver that may itself hold up to 2^30 cells.
"""
import ctypes
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")


@pytest.fixture(scope="module")
def fsa():
    import floodsim_amd
    floodsim_amd.lib()
    return floodsim_amd


def _small(fsa, **kw):
    nx, ny = 64, 48
    g = fsa.GPUSolver(nx=nx, ny=ny, dx=1.0, z=np.zeros(nx * ny), manning=0.03, **kw)
    g.setDepth(lambda X, Y, Z: np.where(X < 20, 1.0, 0.0))
    return g


def test_stall_reports_no_progress(fsa):
    from floodsim_amd.solver import FloodSimError
    g = _small(fsa)
    g._debug_no_steps(True)
    with pytest.raises(FloodSimError) as e:
        g.advance(5.0)
    assert "-4" in str(e.value) and "GPU solver made no progress (device error?)" in str(e.value)
    assert g.t == 0.0 and g.steps == 0
    g._debug_no_steps(False)
    c = g.advance(5.0)
    assert c.t == 5.0 and c.steps > 0
    g.destroy()


STALL_JS = r"""
const path = require('path');
const { GPUSolver } = require(path.join(process.cwd(), 'floodsim_amd', 'js', 'solver-hip.js'));
(async () => {
  const nx = 64, ny = 48, s = new GPUSolver(null, { nx, ny, dx: 1, z: new Float64Array(nx * ny), manning: 0.03 });
  await s.ready;
  s.setDepth((x) => (x < 20 ? 1 : 0));
  s._debugNoSteps(true);
  let msg = 'resolved';
  try { await s.advance(5); } catch (e) { msg = e.message; }
  s._debugNoSteps(false);
  await s.advance(5);
  process.stdout.write(JSON.stringify({ msg, t: s.t, steps: s.steps }) + '\n');
  s.destroy();
})().catch((e) => { process.stderr.write(String(e && e.stack || e) + '\n'); process.exit(1); });
"""


@pytest.mark.skipif(NODE is None, reason="node is not installed")
def test_stall_js():
    print("\nnode stall case", flush=True)
    r = subprocess.run([NODE, "-e", STALL_JS], cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    o = json.loads(r.stdout)
    assert o["msg"] == "GPU solver made no progress (device error?)"
    assert o["t"] == 5 and o["steps"] > 0


def test_tile_stats_counts(fsa):
    g = _small(fsa)
    g.advance(2.0)
    a = g.tileStats()
    strips = -(-64 // 58) * -(-48 // 64)
    assert a["tiled"] and a["tiles"] == strips
    assert 0 < a["update"] <= a["tiles"] and a["flux"] == a["update"]
    assert 0 < a["active"] <= 64 * 48 and a["cells"] == 64 * 48
    b = _small(fsa, tiles=False)
    b.advance(2.0)
    s = b.tileStats()
    assert not s["tiled"] and s["tiles"] == strips and s["active"] == 64 * 48
    c = _small(fsa, lts=3)
    c.advance(2.0)
    t = c.tileStats()
    assert t["tiled"] and t["tiles"] == 8 * 6 and t["update"] == 8 * 6
    # a wide dry domain: strips far from the water stay quiet
    nx, ny = 640, 256
    w = fsa.GPUSolver(nx=nx, ny=ny, dx=1.0, z=np.zeros(nx * ny), manning=0.03)
    w.setDepth(lambda X, Y, Z: np.where(X < 30, 1.0, 0.0))
    w.advance(3.0)
    u = w.tileStats()
    print(f"\ntileStats: {a}, untiled {s}, lts {t}, wide dry domain {u}")
    assert u["tiled"] and u["update"] < u["tiles"] // 2
    for o in (g, b, c, w):
        o.destroy()


def test_read_inlets_depth_and_clear(fsa):
    nx, ny, dx = 60, 40, 2.0
    I, J = np.meshgrid(np.arange(nx), np.arange(ny))
    z = 0.01 * I + 0.002 * (J - 20) ** 2
    inlets = [5 * nx + 5, 20 * nx + 2, 30 * nx + 40, 12 * nx + 50]
    g = fsa.GPUSolver(nx=nx, ny=ny, dx=dx, z=z, manning=0.02, bc={"west": "outflow"}, inlets=inlets)
    g.setDepth(lambda X, Y, Z: 0 * X)
    g.rainRate = 60 / 1000 / 3600
    g.setInletRates([0.05, 0.05, -0.002, 0.5])
    g.advance(120.0)
    r = g.readInlets()
    s = g.readState()
    # global stepping: the exchange is the step's last change of the inlet cells
    np.testing.assert_array_equal(r["h"], s["h"][inlets].astype(np.float32).astype(np.float64))
    assert r["h"][2] > 0 and r["vol"][2] < 0
    r2 = g.readInlets()
    assert np.all(r2["h"] == 0) and np.all(r2["vol"] == 0)
    g.destroy()


def test_16_byte_planes_limit(fsa):
    """A global-stepping, hazard-off grid of more than 2^28 cells is valid, but its Green-Ampt, GI and barrier
    planes would overflow the 32-bit offsets: each is refused with FS_E_ARG."""
    from floodsim_amd.solver import lib, make_params
    L = lib()
    nx, ny = 16385, 16384
    n = nx * ny
    assert n > 1 << 28
    print(f"\ncreating a {nx} x {ny} global-stepping solver", flush=True)
    p = make_params(nx, ny, 1.0)
    z = np.zeros(n)
    h = ctypes.c_void_p()
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    rc = L.fs_create(ctypes.byref(p), dp(z), None, 0.03, None, None, None, None, 0, ctypes.byref(h))
    assert rc == 0, L.fs_last_error(None).decode()
    try:
        one = np.ones(n)
        assert L.fs_set_ga(h, dp(one), dp(one), dp(one)) == -1
        assert b"2^28" in L.fs_last_error(h)
        assert L.fs_set_gi(h, dp(one), dp(one), dp(one)) == -1
        assert b"2^28" in L.fs_last_error(h)
        ip = ctypes.POINTER(ctypes.c_int32)
        i = np.array([5], dtype=np.int32)
        zz = np.array([1.0])
        assert L.fs_set_crests(h, i.ctypes.data_as(ip), i.ctypes.data_as(ip), np.zeros(1, np.int32).ctypes.data_as(ip),
                               dp(zz), dp(zz), dp(zz), 1) == -1
        assert b"2^28" in L.fs_last_error(h)
    finally:
        L.fs_destroy(h)

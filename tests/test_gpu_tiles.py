"""Dry-strip skipping under global stepping: the same bits as advancing every strip.

The cases are the reference's active-tile checks (tests/gpu-suite.js:204-240): a pond breaking
over a 640 x 480 slope (15.8 % of cells wet after 240 s in the reference's run, max|h_tiled -
h_full| = 0) and rain on all of a 2112 x 2112 grid (every tile active, the flow kept). Both run
here with tiles on and off; the state and the books must agree exactly, and the recorded
aggregates of the reference must hold.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MMHR = 1 / 1000 / 3600


def _run(opts, init, rain, frames, tiles):
    import floodsim_amd as fsa
    g = fsa.GPUSolver(**opts, tiles=tiles)
    g.setDepth(init)
    g.rainRate = rain
    stats = []
    for t in frames:
        g.advance(t)
        stats.append(g.tileStats()["active"])
    s = g.readState()
    steps = g.steps
    g.destroy()
    return s, steps, stats


def _same(a, b):
    for k in ("h", "hu", "hv", "hmax"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("volume", "rain", "infil", "outflow", "clipped"):
        assert a["mass"][k] == b["mass"][k], k


def test_pond_break_tiled_equals_untiled(golden):
    nx, ny, dx = 640, 480, 2.0
    j, i = np.mgrid[0:ny, 0:nx].astype(np.float64)
    z = 50 - 0.004 * i * dx + 0.6 * np.sin(j / 23) * np.cos(i / 31) + 0.002 * np.abs(j - ny / 2) * dx
    opts = dict(nx=nx, ny=ny, dx=dx, z=z.reshape(-1), manning=0.03, bc={"east": "outflow"})
    init = lambda X, Y, Z: np.where(np.hypot(X - 120, Y - 480) < 60, np.maximum(0, 53 - Z), 0.0)
    frames = [1.0, 60.0, 240.0]
    a, sa, _ = _run(opts, init, 0.0, frames, tiles=False)
    b, sb, act = _run(opts, init, 0.0, frames, tiles=True)
    assert sa == sb
    _same(a, b)
    wet = float(np.mean(b["h"] > 1e-3))
    ref = golden["gpu"]["tiles"]
    print(f"pond: {sb} steps, active open cells per frame {act} of {nx * ny}, wet {wet:.3f} "
          f"(reference {ref['wetFrac']:.3f}), mass rel {b['mass']['rel']:.1e}")
    assert abs(wet - ref["wetFrac"]) < 0.02
    assert act[0] < 0.25 * nx * ny and act[-1] < nx * ny
    assert abs(b["mass"]["rel"]) <= 1e-6


def test_rain_everywhere_keeps_every_strip(golden):
    nx = ny = 2112
    dx = 2.0
    j, i = np.mgrid[0:ny, 0:nx].astype(np.float64)
    z = 30 - 0.002 * i * dx + 0.4 * np.sin(j / 37) * np.cos(i / 29)
    opts = dict(nx=nx, ny=ny, dx=dx, z=z.reshape(-1), manning=0.03, bc={"east": "outflow"})
    a, sa, _ = _run(opts, lambda X, Y, Z: 0 * X, 80 * MMHR, [300.0], tiles=False)
    b, sb, act = _run(opts, lambda X, Y, Z: 0 * X, 80 * MMHR, [300.0], tiles=True)
    assert sa == sb
    _same(a, b)
    moving = float(np.mean(np.hypot(b["hu"], b["hv"]) > 1e-6))
    print(f"rain on {nx}x{ny}: active {act[-1]} of {nx * ny}, moving {moving:.5f} "
          f"(reference {golden['gpu']['manyTiles']['movingFrac']:.5f}), mass rel {b['mass']['rel']:.1e}")
    assert act[-1] == nx * ny
    assert moving > 0.99
    assert abs(b["mass"]["rel"]) <= 1e-6


def test_inflow_into_dry_domain_and_first_order():
    """A point inflow wakes quiet strips through the exchange; first order takes the same path."""
    nx, ny = 300, 260
    x = (np.arange(nx) + 0.5)
    z = np.tile(0.002 * (nx - x), (ny, 1)).reshape(-1)
    for so in (True, False):
        opts = dict(nx=nx, ny=ny, dx=1.0, z=z, manning=0.03, bc={"east": "outflow"},
                    inlets=[130 * nx + 2], secondOrder=so)
        res = []
        for tiles in (False, True):
            import floodsim_amd as fsa
            g = fsa.GPUSolver(**opts, tiles=tiles)
            g.setDepth(lambda X, Y, Z: 0 * X)
            g.setInletRates([-10.0])
            acts = []
            for t in (5.0, 120.0, 900.0):
                g.advance(t)
                acts.append(g.tileStats()["active"])
            res.append((g.readState(), g.steps, acts))
            g.destroy()
        (a, sa, _), (b, sb, acts) = res
        assert sa == sb
        _same(a, b)
        print(f"inflow secondOrder={so}: {sb} steps, active cells {acts} of {nx * ny}")
        assert 0 < acts[0] < acts[-1] <= nx * ny
        assert abs(b["mass"]["rel"]) <= 1e-6

"""GPU local time stepping against the FP64 oracle's LTS (a C restatement of solver-cpu.js
_cycle/_record, itself pinned by tests/test_oracle.py). Cases follow tests/run-lts-tests.mjs and the
reference GPU suite's LTS section (results-gpu.json `lts`: cycles, vsCPU 2.9e-7 m on the dam breaks,
3.9e-5 m on the 256x192 channel)."""
import numpy as np
import pytest

import oracle
from oracle import refcases as rc

pytestmark = pytest.mark.gpu
TOL_H = 1e-3
TOL_MASS = 1e-6


@pytest.fixture(scope="module")
def fsa():
    import floodsim_amd
    floodsim_amd.lib()
    return floodsim_amd


def pair(fsa, opts, init, tEnd, rain=0.0, lts=3, frame=None, **gopts):
    g = fsa.GPUSolver(**opts, lts=lts, **gopts)
    c = oracle.CPUSolver(**opts, lts=lts, hdry=g.hdry)
    init(g)
    init(c)
    if rain:
        g.rainRate = rain
        c.rainRate = rain
    # frames cut the cycles as the reference app does (10 s quanta, app/map.js:2281)
    for t in (np.arange(frame, tEnd + 1e-9, frame) if frame else [tEnd]):
        g.advance(float(t), 8)
        c.run(float(t))
    s = g.readState()
    d = float(np.max(np.abs(s["h"] - c.field("h"))))
    clk = g.clock()
    print(f"\nLTS: max|dh| {d:.2e} m, cycles gpu/cpu {clk.steps}/{c.steps}, cell updates gpu/cpu "
          f"{clk.cellUpdates}/{c.cellSteps}, mass rel {s['mass']['rel']:.2e}")
    return g, c, s, d


@pytest.mark.parametrize("name,hR", [("stoker", 0.5), ("ritter", 0.0)])
def test_lts_dam_break(fsa, golden, name, hR):
    nx = 400
    opts = dict(nx=nx, ny=3, dx=100 / nx, z=np.zeros(nx * 3), manning=0)
    g, c, s, d = pair(fsa, opts, lambda o: o.setDepth(lambda X, Y, Z: np.where(X < 50, 1.0, hR)), 6)
    assert d <= TOL_H
    assert abs(s["mass"]["rel"]) <= TOL_MASS
    assert g.steps == c.steps == golden["gpu"]["lts"][name]["cycles"]
    assert g.clock().cellUpdates == c.cellSteps


def channel_opts(still=False):
    nx, ny = 256, 192
    I, J = np.meshgrid(np.arange(nx), np.arange(ny))
    z = np.where((J >= 90) & (J < 102), -7.0, 0.3 - 0.002 * I + (0 if still else 0.01 * np.sin(J / 5)))
    return dict(nx=nx, ny=ny, dx=1, z=z, manning=0.03, bc={} if still else {"east": "outflow"})


def test_lts_still_water(fsa):
    """run-lts-tests.mjs:51-58: still water over a deep channel beside a shallow floodplain stays still."""
    g = fsa.GPUSolver(**channel_opts(True), lts=3)
    g.setDepth(lambda X, Y, Z: np.maximum(0, 0.1 - Z))
    g.advance(20, 8)
    s = g.readState()
    w = s["h"] > 1e-5
    vmax = float(np.max(np.hypot(s["hu"][w], s["hv"][w]) / s["h"][w]))
    print(f"\nstill water LTS: max|u| {vmax:.2e} m/s, mass rel {s['mass']['rel']:.2e}, cycles {g.steps}")
    assert vmax <= 1e-4 and abs(s["mass"]["rel"]) <= TOL_MASS


def test_lts_channel(fsa, golden):
    """run-lts-tests.mjs:59-76 on the GPU: LTS against the oracle's LTS, 60 s."""
    opts = channel_opts()
    g, c, s, d = pair(fsa, opts, lambda o: o.setDepth(
        lambda X, Y, Z: np.where(Z < -1, np.where(X < 40, 0.5, 0) - Z, 0.05)), 60)
    r = golden["gpu"]["lts"]["channel256"]
    assert d <= TOL_H and abs(s["mass"]["rel"]) <= TOL_MASS
    assert g.steps == c.steps == r["cycles"]
    assert 2095 * 256 * 192 / g.clock().cellUpdates >= 2.5


def test_lts_urban_rain(fsa):
    """The urban cloudburst (src/scenarios.js:8-49) under 150 mm/h with 4 LTS levels, 600 s in 10 s
    frames. (One 600 s advance from a dry start is a degenerate cycle: its length comes from the
    speed of the first millimetres of rain, and the film outgrows that step within the cycle.)"""
    sc = rc.cloudburst(80)
    opts = dict(nx=sc["nx"], ny=sc["ny"], dx=sc["dx"], z=sc["z"], mask=sc["mask"], manning=sc["manning"],
                bc=sc["bc"], infil=sc["infil"])
    g, c, s, d = pair(fsa, opts, lambda o: o.setDepth(sc["depth"]), 600, rain=150 * rc.MMHR, lts=4, frame=10)
    assert d <= TOL_H and abs(s["mass"]["rel"]) <= TOL_MASS and s["mass"]["clipped"] == 0
    assert abs(s["mass"]["outflow"] - c.vol["outflow"]) <= 1e-3 * c.vol["outflow"] + 1e-6


def test_lts_record_parts_change_nothing(fsa):
    """A record walks the runs of rows its units own, cut into parts of at most 32 rows, or, for a record of
    few owned rows, of 1, 2, 4, 8 or 16 rows (within 4096 wavefronts). The unit height (8-40 rows here
    against the default 64) changes every run and part; with 40-row units the last unit is cut by the
    domain's edge. Every cell is updated by one part with the same arithmetic, so the state is the 64-row
    run's bit for bit (and the oracle's to the parity bar), with the books closed."""
    sc = rc.cloudburst(80)
    opts = dict(nx=sc["nx"], ny=sc["ny"], dx=sc["dx"], z=sc["z"], mask=sc["mask"], manning=sc["manning"],
                bc=sc["bc"], infil=sc["infil"])
    init = lambda o: o.setDepth(sc["depth"])
    _, c, ref, d = pair(fsa, opts, init, 120, rain=150 * rc.MMHR, lts=4, frame=10)
    assert d <= TOL_H
    for rows in (8, 16, 24, 40):
        g = fsa.GPUSolver(**opts, lts=4, rows=rows)
        init(g)
        g.rainRate = 150 * rc.MMHR
        for t in np.arange(10, 120 + 1e-9, 10):
            g.advance(float(t), 8)
        s = g.readState()
        dr = max(float(np.max(np.abs(s[k] - ref[k]))) for k in ("h", "hu", "hv", "hmax"))
        same = all(np.array_equal(s[k], ref[k]) for k in ("h", "hu", "hv", "hmax"))
        print(f"  rows {rows}: max |d| against 64 rows {dr:.2e} (bit-identical {same}), against the oracle "
              f"{float(np.max(np.abs(s['h'] - c.field('h')))):.2e} m, mass rel {s['mass']['rel']:.1e}")
        assert same and dr == 0.0
        assert abs(s["mass"]["rel"]) <= TOL_MASS and s["mass"]["clipped"] == 0
        g.destroy()


def _varied_case(seed, nx, ny):
    """A bumpy basin with a deep pool in one corner, a shallow film elsewhere and building walls: the pool's
    fast water puts its tiles on the fine levels, the film on the coarse ones, with walls at every level."""
    rng = np.random.default_rng(seed)
    I, J = np.meshgrid(np.arange(nx), np.arange(ny))
    z = 0.004 * I + 0.002 * J + 0.3 * np.sin(I / 7.0 + seed) * np.cos(J / 5.0)
    mask = np.ones((ny, nx), np.uint8)
    for _ in range(12):
        w, h = rng.integers(3, 12, 2)
        x0, y0 = rng.integers(0, nx - w), rng.integers(0, ny - h)
        mask[y0:y0 + h, x0:x0 + w] = 0
    mask[: ny // 4, : nx // 4] = 1    # the pool stays open
    opts = dict(nx=nx, ny=ny, dx=1.0, z=z.ravel(), mask=mask.ravel(), manning=0.02,
                bc={"east": "outflow", "north": "outflow"})
    def init(o):
        o.setDepth(lambda X, Y, Z: np.where((X < nx / 4) & (Y < ny / 4), 1.5, 0.02))
    return opts, init


@pytest.mark.parametrize("seed,nx,ny", [(1, 150, 110), (2, 97, 203)])
def test_lts_varied_levels(fsa, seed, nx, ny):
    """Levels 0-3 occupied in patterns that change from frame to frame (the oracle's levels at 5, 10, 20 and 30 s:
    seed 1 [0, 45, 16, 205, 0], [46, 42, 178, 0, 0], [0, 0, 129, 137, 0], [99, 150, 17, 0, 0] tiles per level),
    on grids that are multiples of neither the 8x8 tile nor the 56-column unit: against the oracle's LTS
    (cycles, cell updates, depths, books) and with 8-row units against the default units bit for bit (every
    run and part different)."""
    opts, init = _varied_case(seed, nx, ny)
    g, c, s, d = pair(fsa, opts, init, 40, rain=60 * rc.MMHR, lts=4, frame=10)
    assert d <= TOL_H and abs(s["mass"]["rel"]) <= TOL_MASS
    assert g.steps == c.steps and g.clock().cellUpdates == c.cellSteps
    assert abs(s["mass"]["outflow"] - c.vol["outflow"]) <= 1e-3 * max(c.vol["outflow"], 1.0)
    h = fsa.GPUSolver(**opts, lts=4, rows=8)
    init(h)
    h.rainRate = 60 * rc.MMHR
    for t in np.arange(10, 40 + 1e-9, 10):
        h.advance(float(t), 8)
    t8 = h.readState()
    for k in ("h", "hu", "hv", "hmax"):
        np.testing.assert_array_equal(t8[k], s[k])
    h.destroy()
    g.destroy()


@pytest.mark.parametrize("second_order", [True, False])
def test_lts_pool_and_film(fsa, second_order):
    """A deep pool spreading over a conical slope wetted by a 0.8 mm film, east edge free: the pool's fast water
    takes the fine levels and the film the coarse ones, so a waiting film cell beside a finer one passes water
    out (the draining-time limiter with its store) and every level from 0 to 4 is occupied in some frame (the
    oracle's tiles per level at 5 and 20 s: [68, 22, 12, 12, 30] and [0, 0, 0, 20, 124]). Both orders: first
    order adds a record's gathers after it (sw_lts_fold), second order in its second stage. Against the
    oracle's LTS: cycles and cell updates equal, depths (1.3e-6 / 1.4e-7 m), books."""
    nx = ny = 96
    I, J = np.meshgrid(np.arange(nx), np.arange(ny))
    z = 0.02 * np.hypot(I - 20, J - 48) / 10
    opts = dict(nx=nx, ny=ny, dx=1.0, z=z.ravel(), manning=0.03, bc={"east": "outflow"}, secondOrder=second_order)
    init = lambda o: o.setDepth(lambda X, Y, Z: np.where(np.hypot(X - 20, Y - 48) < 12, 2.0, 0.0008))
    g, c, s, d = pair(fsa, opts, init, 30, lts=4, frame=5)
    assert d <= TOL_H and abs(s["mass"]["rel"]) <= TOL_MASS
    assert g.steps == c.steps and g.clock().cellUpdates == c.cellSteps
    assert abs(s["mass"]["outflow"] - c.vol["outflow"]) <= 1e-3 * max(c.vol["outflow"], 1.0)


@pytest.mark.parametrize("second_order", [True, False])
def test_lts_inlets_pool_and_film(fsa, second_order):
    """The pool-and-film case with three drainage inlets: a sink in the fast pool (a fine level), a point inflow
    and a sink on the slow film (coarse levels). Each inlet exchanges at the end of its cell's level records
    (exchangeL, kernels.wgsl.js:596), at both orders (first order: after the record's gather walk). Against the
    oracle's LTS: cycles, depths, each inlet's exchanged volume and the books."""
    nx = ny = 96
    I, J = np.meshgrid(np.arange(nx), np.arange(ny))
    z = 0.02 * np.hypot(I - 20, J - 48) / 10
    inlets = [48 * nx + 20, 48 * nx + 60, 20 * nx + 40]
    opts = dict(nx=nx, ny=ny, dx=1.0, z=z.ravel(), manning=0.03, bc={"east": "outflow"}, secondOrder=second_order,
                inlets=inlets)

    def init(o):
        o.setDepth(lambda X, Y, Z: np.where(np.hypot(X - 20, Y - 48) < 12, 2.0, 0.0008))
        o.setInletRates([0.5, -0.2, 0.05])
    g, c, s, d = pair(fsa, opts, init, 30, lts=4, frame=5)
    vg, vc = g.readInlets()["vol"], c.readInlets()["vol"]
    print(f"  inlet volumes gpu {np.round(vg, 4).tolist()} oracle {np.round(vc, 4).tolist()} m3")
    assert d <= TOL_H and abs(s["mass"]["rel"]) <= TOL_MASS
    assert g.steps == c.steps and g.clock().cellUpdates == c.cellSteps
    assert np.all(np.abs(vg - vc) <= 1e-3 * np.abs(vc) + 1e-6)
    assert abs(s["mass"]["exchanged"] - c.vol["exchanged"]) <= 1e-3 * abs(c.vol["exchanged"]) + 1e-6

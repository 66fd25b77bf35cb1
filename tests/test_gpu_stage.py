"""GPU water-level boundary (stageCells + setStage) against the FP64 oracle, itself pinned by
tests/test_oracle.py to the reference's outlet known answers (tests/run-outlet-tests.mjs) and to the
inflow the reference GPU recorded for its stage case (results-gpu.json `stage`).

Bars: max |dh| <= 1 mm and |mass rel| <= 1e-6, as for every parity case."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
TOL_H = 1e-3
TOL_MASS = 1e-6


@pytest.fixture(scope="module")
def fsa():
    import floodsim_amd
    floodsim_amd.lib()
    return floodsim_amd


def sea(lts=0):
    nx, ny, dx = 50, 20, 4
    I, J = np.meshgrid(np.arange(nx), np.arange(ny))
    z = (98 + 0.08 * I + 0.3 * np.sin(J)).reshape(-1)
    return dict(nx=nx, ny=ny, dx=dx, z=z, manning=0.02, bc={k: "outflow" for k in ("west", "east", "south", "north")},
                stageCells=[j * nx for j in range(ny)], lts=lts)


def drive_sea(s, z0=99.5, start_wet=True):
    """gpu-suite.js:148-164: still water at the stage for 300 s, then the stage rises 0.8 m over 900 s
    in 30 s frames."""
    s.setStage(z0)
    s.setDepth(lambda X, Y, Z: np.maximum(0, z0 - Z) if start_wet else 0 * X)
    yield 300.0
    for t in range(330, 1801, 30):
        s.setStage(z0 + 0.8 * min(1, (t - 300) / 900))
        yield float(t)


def run_gpu(fsa, opts, start_wet=True, K=64):
    g = fsa.GPUSolver(**opts)
    rest = None
    for t in drive_sea(g, start_wet=start_wet):
        g.advance(t, K)
        if t == 300.0:
            s = g.readState()
            wet = s["h"] > 1e-3
            umax = float(np.max(np.hypot(s["hu"][wet], s["hv"][wet]) / s["h"][wet])) if wet.any() else 0.0
            rest = (umax, s["mass"]["outflow"])
    return g, g.readState(), rest


def run_oracle(opts, start_wet=True):
    c = oracle.CPUSolver(**opts, hdry=1e-5)
    for t in drive_sea(c, start_wet=start_wet):
        c.run(t)
    return c


def test_stage_sea(fsa, golden):
    """Still water at the stage stays still (the reference GPU: max|u| 5.8e-6 m/s, boundary flux 1.8e-6 m^3);
    the rising stage floods in through the boundary: 5699.31 m^3 on the reference GPU."""
    opts = sea()
    g, s, rest = run_gpu(fsa, opts)
    c = run_oracle(opts)
    d = float(np.max(np.abs(s["h"] - c.field("h"))))
    ref = golden["gpu"]["stage"]
    inflow = -s["mass"]["outflow"]
    print(f"\nstage sea: max|dh| {d:.2e} m, at rest max|u| {rest[0]:.1e} m/s (reference {ref['restUmax']:.1e}), "
          f"boundary flux {rest[1]:.1e} m3; inflow {inflow:.2f} m3 (oracle {-c.vol['outflow']:.2f}, reference GPU "
          f"{ref['inflow']:.2f}), mass rel {s['mass']['rel']:.1e}")
    assert d <= TOL_H and abs(s["mass"]["rel"]) <= TOL_MASS
    assert rest[0] <= 1e-4 and abs(rest[1]) <= 1e-3
    assert abs(inflow - ref["inflow"]) <= 1e-3 * ref["inflow"]
    assert abs(inflow + c.vol["outflow"]) <= 1e-3 * inflow


def test_stage_sea_dry_start_tiles(fsa):
    """A stage rising onto a dry shore: the strips holding stage cells always step (the reference keeps its
    stage-edge tiles active, solver-gpu.js:134-143), so skipping quiet strips gives the same bits as
    advancing every strip, and the oracle's answer."""
    opts = sea()
    _, a, _ = run_gpu(fsa, opts, start_wet=False)
    _, b, _ = run_gpu(fsa, dict(opts, tiles=False), start_wet=False)
    for k in ("h", "hu", "hv", "hmax"):
        np.testing.assert_array_equal(a[k], b[k])
    c = run_oracle(opts, start_wet=False)
    d = float(np.max(np.abs(a["h"] - c.field("h"))))
    print(f"\nstage, dry start: max|dh| {d:.2e} m, inflow {-a['mass']['outflow']:.2f} m3")
    assert d <= TOL_H and abs(a["mass"]["rel"]) <= TOL_MASS and -a["mass"]["outflow"] > 1000


def test_stage_sea_lts(fsa):
    opts = sea(lts=3)
    g, s, _ = run_gpu(fsa, opts, K=8)
    c = run_oracle(opts)
    d = float(np.max(np.abs(s["h"] - c.field("h"))))
    print(f"\nstage sea, LTS: max|dh| {d:.2e} m, cycles {g.steps}/{c.steps}, mass rel {s['mass']['rel']:.1e}")
    assert d <= TOL_H and abs(s["mass"]["rel"]) <= TOL_MASS


def test_outlet_channel_cut(fsa):
    """run-outlet-tests.mjs:35-55: the cut river model, 200 m^3/s in at its west end (inflow inlets) and its
    east edge held at the backwater stage, for 4000 s."""
    W, dx, n, Q, down, nx = 5, 10.0, 0.03, 200.0, 3.0, 80
    edge = (down ** (13 / 3) + 13 / 3 * (Q / (W * dx) * n) ** 2 * 800) ** (3 / 13)
    opts = dict(nx=nx, ny=W, dx=dx, z=np.zeros(nx * W), manning=n,
                bc={"west": "wall", "south": "wall", "north": "wall", "east": "outflow"},
                stageCells=[j * nx + nx - 1 for j in range(W)], inlets=[j * nx for j in range(W)])
    g = fsa.GPUSolver(**opts)
    c = oracle.CPUSolver(**opts, hdry=g.hdry)
    for o in (g, c):
        o.setStage(edge)
        o.setDepth(lambda X, Y, Z: edge + 0 * X)
        o.setInletRates(np.full(W, -Q / W))
    g.advance(4000)
    c.run(4000)
    s = g.readState()
    d = float(np.max(np.abs(s["h"] - c.field("h"))))
    ws20 = s["h"].reshape(W, nx)[:, 20].mean()
    print(f"\noutlet cut model: max|dh| {d:.2e} m, level at 200 m {ws20:.4f} m, steps {g.steps}/{c.steps}, "
          f"mass rel {s['mass']['rel']:.1e}")
    assert d <= TOL_H and abs(s["mass"]["rel"]) <= TOL_MASS

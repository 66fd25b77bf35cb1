"""fs_advance_noalloc: fs_advance with its descriptor, work list and per-wavefront partials in the library's
fixed device array, no allocation. Checked through the C ABI against fs_advance on the same inputs, bit for
bit (the same kernels in the same order): three calls in a row, each on a smaller solver than the one before
(the array keeps the last call's contents), under local time stepping with an inflow, global stepping with
rain, and first order; then the errors: a grid whose work list does not fit the array, and timing mode on."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("h", "hu", "hv", "hmax")


def _solver(fsa, case):
    g = fsa.GPUSolver(**case["opts"])
    g.setDepth(case["depth"])
    g.rainRate = case["rain"]
    if case.get("rates") is not None:
        g.setInletRates(case["rates"])
    return g


def _cases():
    from floodsim_amd import scenarios as sc
    u = sc.urban(384, seed=3)
    big = dict(name="urban 384^2, LTS 4 levels, rain and a point inflow",
               opts=dict(nx=384, ny=384, dx=1.0, z=u["z"], mask=u["mask"], manning=u["manning"],
                         bc={k: "outflow" for k in ("west", "east", "south", "north")}, inlets=[192 * 384], lts=4,
                         hazard=True),
               depth=lambda X, Y, Z: 0 * X, rain=50 * sc.MMHR, rates=[-2.0], frames=(60.0, 120.0, 180.0))
    v = sc.urban(200, seed=4, ny=150)
    mid = dict(name="urban 200x150, global steps, rain",
               opts=dict(nx=200, ny=150, dx=1.0, z=v["z"], mask=v["mask"], manning=v["manning"],
                         bc={"east": "outflow"}),
               depth=lambda X, Y, Z: 0 * X, rain=80 * sc.MMHR, frames=(60.0, 150.0))
    small = dict(name="dam break 96x24, first order",
                 opts=dict(nx=96, ny=24, dx=1.0, z=np.zeros(96 * 24), manning=0.03, secondOrder=False),
                 depth=lambda X, Y, Z: np.where(X < 40, 1.0, 0.1), rain=0.0, frames=(5.0, 10.0))
    return [big, mid, small]


def test_noalloc_equals_advance():
    import floodsim_amd as fsa
    from floodsim_amd.solver import fs_clock, lib
    L = lib()
    for case in _cases():
        print(f"\nno-alloc advance: {case['name']}", flush=True)
        ref = _solver(fsa, case)
        new = _solver(fsa, case)
        for t in case["frames"]:
            cr = ref.advance(t)
            cn = fs_clock()
            rc = L.fs_advance_noalloc(new._h, float(t), 64, ctypes.byref(cn))
            assert rc == 0, L.fs_last_error(new._h).decode()
            assert (cn.t, cn.steps, cn.lastDt, cn.cellUpdates) == (cr.t, cr.steps, cr.lastDt, cr.cellUpdates)
        a, b = ref.readState(), new.readState()
        for k in FIELDS:
            np.testing.assert_array_equal(a[k], b[k])
        assert a["mass"] == b["mass"]
        print(f"  {cr.steps} steps to t = {cr.t}: equal bit for bit, mass rel {b['mass']['rel']:.1e}")
        assert abs(b["mass"]["rel"]) <= 1e-6
        ref.destroy()
        new.destroy()


def test_noalloc_errors():
    import floodsim_amd as fsa
    from floodsim_amd.solver import fs_clock, lib
    L = lib()
    nx, ny = 58 * 256, 8 * 256     # 65536 strips of 58 x 8 cells: a work list beyond the fixed array
    print(f"\nno-alloc advance: {nx} x {ny} at 8 rows per strip must be refused", flush=True)
    g = fsa.GPUSolver(nx=nx, ny=ny, dx=1.0, z=np.zeros(nx * ny), manning=0.03, rows=8)
    c = fs_clock()
    assert L.fs_advance_noalloc(g._h, 1.0, 4, ctypes.byref(c)) == -1
    assert b"scratch" in L.fs_last_error(g._h)
    assert g.clock().steps == 0
    g.destroy()
    # the header's bound: 8192 x 8192 fits with local time stepping too (its records hold three partials a unit)
    n = 8192
    print(f"no-alloc advance: {n} x {n} with local time stepping must fit", flush=True)
    big = fsa.GPUSolver(nx=n, ny=n, dx=1.0, z=np.zeros(n * n), manning=0.03, lts=4)
    big.setDepth(lambda X, Y, Z: np.where(X < 64, 0.5, 0.0))
    assert L.fs_advance_noalloc(big._h, 0.5, 4, ctypes.byref(c)) == 0 and c.t == 0.5
    big.destroy()
    s = fsa.GPUSolver(nx=64, ny=32, dx=1.0, z=np.zeros(64 * 32), manning=0.03)
    s.setDepth(lambda X, Y, Z: np.where(X < 20, 1.0, 0.0))
    s.set_timing(True)
    assert L.fs_advance_noalloc(s._h, 1.0, 4, ctypes.byref(c)) == -5
    s.set_timing(False)
    assert L.fs_advance_noalloc(s._h, 1.0, 4, ctypes.byref(c)) == 0 and c.t == 1.0
    s.destroy()

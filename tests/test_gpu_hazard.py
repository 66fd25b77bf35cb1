"""Hazard products every step on the GPU against the FP64 oracle: the reference's own check
(tests/gpu-suite.js:264-282, a dam break over a bumpy slope, GPU vs CPU tracker) under global
and local time stepping. The reference recorded 0 class mismatches, force within 1.1e-4 kN/m,
d*v within 2e-5 m^2/s, duration within 0.27 s, arrival within 4e-6 s (results-gpu.json hazard)."""
import math

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _case():
    nx, ny, dx = 80, 40, 2.0
    j, i = np.mgrid[0:ny, 0:nx].astype(np.float64)
    z = 3 - 0.03 * i + 0.2 * np.sin(i / 5) * np.cos(j / 4)
    opts = dict(nx=nx, ny=ny, dx=dx, z=z.reshape(-1), manning=0.035,
                bc={"west": "wall", "east": "outflow", "south": "wall", "north": "wall"})
    init = lambda X, Y, Z: np.where(X < 30, np.maximum(0, 4.5 - Z), 0.0)
    return opts, init


@pytest.mark.parametrize("lts", [0, 3])
def test_hazard_vs_oracle(golden, lts):
    import floodsim_amd as fsa
    opts, init = _case()
    g = fsa.GPUSolver(**opts, hazard=True, lts=lts)
    c = oracle.CPUSolver(**opts, hdry=g.hdry, lts=lts)
    g.setDepth(init)
    c.setDepth(init)
    for t in (30.0, 120.0, 240.0):
        g.advance(t)
        c.run(t)
    G, C = g.readHazard(), c.readHazard()
    dh = float(np.max(np.abs(g.readState()["h"] - c.field("h"))))
    wet = int(np.sum((C["arr"] > 0) | (G["cls"]["arr"] > 0)))
    cls = sum(int(np.sum(C[k] != G["cls"][k])) for k in ("arr", "defra", "defraUrban"))
    dF = float(np.max(np.abs(C["force"] - G["force"])))
    dDV = float(np.max(np.abs(C["dv"] - G["dv"])))
    dDur = float(np.max(np.abs(C["duration"] - G["duration"])))
    both = ~np.isnan(C["arrival"]) | ~np.isnan(G["arrival"])
    dArr = float(np.max(np.abs(C["arrival"][both] - G["arrival"][both]))) if both.any() else 0.0
    dV = float(np.max(np.abs(C["vmax"] - G["vmax"])))
    miss = int(np.sum(np.isnan(C["tWet"]) != np.isnan(G["tWet"])))
    ref = golden["gpu"]["hazard"]
    print(f"hazard lts={lts}: {wet} cells (reference {ref['wetCells']}), class mismatches {cls}, max|dforce| {dF:.1e} "
          f"kN/m, max|d(dv)| {dDV:.1e}, max|dduration| {dDur:.2f} s, max|darrival| {dArr:.2f} s, max|dvmax| {dV:.1e}, "
          f"first-wet mismatches {miss}; max|dh| {dh:.1e} m")
    assert wet > 0.5 * opts["nx"] * opts["ny"]
    assert cls <= 0.005 * 3 * wet
    assert dF <= 1e-2 and dDV <= 1e-3 and dV <= 1e-2
    assert dDur <= 1.0 and not math.isnan(dArr) and dArr <= 1.0
    assert miss == 0
    g.destroy()


def test_hazard_needs_the_option():
    import floodsim_amd as fsa
    opts, init = _case()
    g = fsa.GPUSolver(**opts)
    with pytest.raises(fsa.FloodSimError):
        g.readHazard()
    g.destroy()

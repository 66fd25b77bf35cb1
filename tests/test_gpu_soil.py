"""GPU Green-Ampt soil and GI stores against the FP64 oracle, itself pinned by tests/test_oracle.py to
the reference CPU's recorded volumes for the same cases (results-gpu.json `gi.cpuStored`,
`greenAmptScour.infilCpu`).

The reference GPU recorded for these cases (tests/gpu-suite.js:188-200, 332-356): GI max |h_gpu - h_cpu|
8.4e-8 m and stored 94.6345 m^3 (CPU 94.6346); Green-Ampt infiltrated within 1.2e-7 of the CPU, max |dh|
1.9e-4 m. Bars here: max |dh| <= 1 mm, |mass rel| <= 1e-6, the volumes within 1e-4 (relative) of the
oracle's and of the reference CPU's."""
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


def both(fsa, opts, sched, **extra):
    g = fsa.GPUSolver(**opts, **extra)
    c = oracle.CPUSolver(**opts, hdry=g.hdry, **extra)
    for o in (g, c):
        o.setDepth(lambda X, Y, Z: 0 * X)
    for rate, t in sched:
        g.rainRate = c.rainRate = rate
        g.advance(t, 8 if extra.get("lts") else 64)
        c.run(t)
    s = g.readState()
    return g, c, s, float(np.max(np.abs(s["h"] - c.field("h"))))


@pytest.mark.parametrize("extra", [{}, dict(lts=3), dict(secondOrder=False)])
def test_gi_stores(fsa, golden, extra):
    opts, sched = rc.gi_case()
    g, c, s, d = both(fsa, opts, sched, **extra)
    store = c.field("giStore").sum() * opts["dx"] ** 2
    gs = s["mass"]["giStored"]
    print(f"\nGI {extra}: max|dh| {d:.2e} m, stored {gs:.4f} m3 (oracle {store:.4f}, reference CPU "
          f"{golden['gpu']['gi']['cpuStored']:.4f}), infiltrated {s['mass']['infil']:.3f}/{c.vol['infil']:.3f}, "
          f"mass rel {s['mass']['rel']:.1e}")
    assert d <= TOL_H and abs(s["mass"]["rel"]) <= TOL_MASS
    assert abs(gs - store) <= 1e-4 * store
    if not extra:
        assert abs(gs - golden["gpu"]["gi"]["cpuStored"]) <= 1e-4 * gs


@pytest.mark.parametrize("extra", [{}, dict(lts=3), dict(secondOrder=False)])
def test_green_ampt(fsa, golden, extra):
    opts, sched = rc.ga_case()
    g, c, s, d = both(fsa, opts, sched, **extra)
    inf = s["mass"]["infil"]
    print(f"\nGreen-Ampt {extra}: max|dh| {d:.2e} m, infiltrated {inf:.4f} m3 (oracle {c.vol['infil']:.4f}, "
          f"reference CPU {golden['gpu']['greenAmptScour']['infilCpu']:.4f}), mass rel {s['mass']['rel']:.1e}")
    assert d <= TOL_H and abs(s["mass"]["rel"]) <= TOL_MASS
    assert abs(inf - c.vol["infil"]) <= 1e-4 * c.vol["infil"]
    if not extra:
        assert abs(inf - golden["gpu"]["greenAmptScour"]["infilCpu"]) <= 1e-4 * inf
    # readInfiltration (solver-gpu.js:507-512): the soil's infiltrated depth per cell, against the oracle's
    F, Fc = g.readInfiltration(), c.field("gaF")
    dF = float(np.max(np.abs(F - Fc)))
    print(f"  readInfiltration: max |dF| {dF:.2e} m, max F {Fc.max():.4f} m")
    assert F.shape == (opts["nx"] * opts["ny"],) and Fc.max() > 0 and dF <= 1e-4 * Fc.max() + 1e-7

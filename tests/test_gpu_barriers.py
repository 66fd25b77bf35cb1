"""GPU flood barriers (crest faces, weir flow, setCrestLevels, barrier books) against the FP64 oracle,
itself pinned by tests/test_oracle.py to the reference CPU's recorded over-volume for the same case
(results-gpu.json `barriers.overVolCPU`).

The case is the reference GPU suite's (tests/gpu-suite.js:242-263): a wall of barrier faces across a
sloping basin and a diagonal leg, overtopped by a dam break, the leg raised 0.5 m at 30 s. The reference
GPU recorded max |h_gpu - h_cpu| 1.9e-5 m, tiled = untiled, books within 5e-7 of the CPU's. Bars here:
max |dh| <= 1 mm, |mass rel| <= 1e-6, tiled = untiled bit for bit, books within 1e-4 (relative) of the
oracle's, levels and loads within 1 mm / 1e-2 kN/m, first overtopping within one step."""
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


def run_gpu(fsa, opts, lift, init, **extra):
    g = fsa.GPUSolver(**opts, **extra)
    g.setDepth(init)
    g.advance(30)
    g.setCrestLevels(lift)
    g.advance(90)
    s = g.readState()
    b = g.readBarriers()
    return g, s, b


def run_oracle(opts, lift, init, **extra):
    c = oracle.CPUSolver(**opts, hdry=1e-5, **extra)
    c.setDepth(init)
    c.run(30)
    c.setCrestLevels(lift)
    c.run(90)
    return c


def compare(name, s, b, c, golden=None):
    cb = c.readBarriers()
    d = float(np.max(np.abs(s["h"] - c.field("h"))))
    over, overc = b["plus"].sum() + b["minus"].sum(), cb["plus"].sum() + cb["minus"].sum()
    fin = np.isfinite(cb["level"])
    dlev = float(np.max(np.abs(cb["level"][fin] - b["level"][fin]))) if fin.any() else 0.0
    dload = float(np.max(np.abs(cb["load"] - b["load"])))
    both = np.isfinite(cb["tOver"]) | np.isfinite(b["tOver"])
    dover = float(np.max(np.abs(cb["tOver"][both] - b["tOver"][both]))) if both.any() else 0.0
    ref = f" (reference GPU {golden['overVol']:.4f}, CPU {golden['overVolCPU']:.4f})" if golden else ""
    print(f"\n{name}: max|dh| {d:.2e} m, over {over:.4f} m3 (oracle {overc:.4f}){ref}, level/load/first-over vs "
          f"oracle {dlev:.1e} m / {dload:.1e} kN/m / {dover:.2f} s, mass rel {s['mass']['rel']:.1e}")
    assert d <= TOL_H and abs(s["mass"]["rel"]) <= TOL_MASS
    assert abs(over - overc) <= 1e-4 * overc
    assert np.array_equal(fin, np.isfinite(b["level"]))
    assert dlev <= 1e-3 and dload <= 1e-2
    return dover


def test_barriers_vs_oracle(fsa, golden):
    opts, lift, init = rc.barrier_case()
    g, s, b = run_gpu(fsa, opts, lift, init, tiles=False)
    _, st, bt = run_gpu(fsa, opts, lift, init, tiles=True)
    c = run_oracle(opts, lift, init)
    dover = compare("barriers", s, b, c, golden["gpu"]["barriers"])
    assert dover <= 1.0   # a step is ~0.2 s here
    for k in ("h", "hu", "hv", "hmax"):
        np.testing.assert_array_equal(s[k], st[k])
    for k in b:
        np.testing.assert_array_equal(b[k], bt[k])
    assert g.steps == c.steps


def test_barriers_lts(fsa):
    """Local time stepping: barrier faces step at the lower level of their two cells, their books are
    kept per record (the reference's barrierL, kernels.wgsl.js:376-379), timed at the cycle start."""
    opts, lift, init = rc.barrier_case()
    _, s, b = run_gpu(fsa, opts, lift, init, lts=3)
    c = run_oracle(dict(opts, lts=3), lift, init)
    compare("barriers, LTS", s, b, c)


def test_crest_below_beds_changes_nothing(fsa):
    """A crest below both beds is no barrier: the same bits as no barrier at all, through the extended
    face kernels."""
    opts, lift, init = rc.barrier_case()
    low = [dict(f, z=-10.0) for f in opts["crest"]]
    runs = []
    for crest in (low, None):
        g = fsa.GPUSolver(**dict(opts, crest=crest))
        g.setDepth(init)
        g.advance(20)
        runs.append(g.readState())
    d = float(np.max(np.abs(runs[0]["h"] - runs[1]["h"])))
    print(f"\ncrest below beds vs none: max|dh| {d:.1e} m")
    assert d <= 1e-6

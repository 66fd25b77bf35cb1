"""Device-state snapshot / restore (the reference's checkpoint of the solver, src/checkpoint.js:128-150):
a run resumed from a snapshot gives the same bits as the uninterrupted run, in the same solver and
in a new solver built from the same model, under global and local time stepping."""
import numpy as np
import pytest

from oracle import refcases as rc

pytestmark = pytest.mark.gpu


def _solver(lts):
    import floodsim_amd as fsa
    sc = rc.cloudburst(80)
    return fsa.GPUSolver(nx=sc["nx"], ny=sc["ny"], dx=sc["dx"], z=sc["z"], mask=sc["mask"], manning=sc["manning"],
                         bc=sc["bc"], lts=lts, inlets=[40 * sc["nx"] + 3])


@pytest.mark.parametrize("lts", [0, 3])
def test_resume_equals_uninterrupted(lts):
    a = _solver(lts)
    a.setDepth(lambda X, Y, Z: 0 * X)
    a.rainRate = 120 * rc.MMHR
    a.setInletRates([-0.5])
    for t in (10.0, 60.0, 120.0):
        a.advance(t)
    snap = a.snapshot()
    c0 = a.clock()
    a.advance(240.0)
    ref = a.readState()
    steps_ref = a.steps
    # the same solver, rewound
    a.restore(snap)
    assert a.clock().t == c0.t and a.clock().steps == c0.steps
    a.advance(240.0)
    again = a.readState()
    # a new solver of the same model
    b = _solver(lts)
    b.restore(snap)
    b.rainRate = 120 * rc.MMHR
    b.advance(240.0)
    fresh = b.readState()
    for s in (again, fresh):
        for k in ("h", "hu", "hv", "hmax"):
            np.testing.assert_array_equal(s[k], ref[k])
        for k in ("volume", "rain", "outflow", "exchanged", "expected"):
            assert s["mass"][k] == ref["mass"][k], k
    assert a.steps == steps_ref and b.steps == steps_ref
    a.destroy()
    b.destroy()


def test_restore_rejects_another_model():
    import floodsim_amd as fsa
    a = _solver(0)
    snap = a.snapshot()
    other = fsa.GPUSolver(nx=40, ny=40, dx=1.0, z=np.zeros(1600), manning=0.03)
    with pytest.raises(fsa.FloodSimError):
        other.restore(snap)
    with pytest.raises(fsa.FloodSimError):
        a.restore(snap[:100])
    a.destroy()
    other.destroy()

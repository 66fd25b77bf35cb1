"""Device-state snapshot / restore (the reference's checkpoint of the solver, src/checkpoint.js:94-113):
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


def _rich_solver(lts, late_field):
    """Every allocation a solver may hold beside its main one: flood barriers (crests, books), Green-Ampt
    soil and GI stores, stage cells, and a rain field given after creation."""
    import floodsim_amd as fsa
    sc = rc.cloudburst(80)
    nx, ny = sc["nx"], sc["ny"]
    n = nx * ny
    z = np.asarray(sc["z"], dtype=np.float64).reshape(-1)
    r = np.random.default_rng(9)
    ga = dict(psiD=np.full(n, 0.05), fmax=np.full(n, 0.3), perv=np.where(r.random(n) < 0.5, 0.6, 0.0))
    gi = dict(rate=np.where(r.random(n) < 0.2, 5e-5, 0.0), cap=np.full(n, 0.05), drain=np.full(n, 1e-6))
    crest = [dict(i=40, j=j, axis=0, z=float(max(z[j * nx + 39], z[j * nx + 40])) + 0.02, C=1.6) for j in range(10, 30)]
    s = fsa.GPUSolver(nx=nx, ny=ny, dx=sc["dx"], z=sc["z"], mask=sc["mask"], manning=sc["manning"], bc=sc["bc"],
                      lts=lts, infil=sc["infil"], stageCells=[j * nx for j in range(ny)], crest=crest, ga=ga, gi=gi)
    s.setStage(float(z.min()) + 0.5)
    if late_field:
        s.setRainField(np.where(np.arange(n) % 4 == 0, 0.0, 1.5))
    return s


@pytest.mark.parametrize("lts", [0, 3])
def test_resume_with_barriers_soil_gi_and_late_rain_field(lts):
    """The barrier books, the soil's infiltrated depth, the GI stores and a rain field given after creation
    live in allocations of their own; a snapshot carries them (the reference saves every storage buffer,
    checkpoint.js:94-103), so a resumed run gives the same bits and books as the uninterrupted one."""
    a = _rich_solver(lts, True)
    a.setDepth(lambda X, Y, Z: 0 * X)
    a.rainRate = 120 * rc.MMHR
    for t in (30.0, 90.0):
        a.advance(t)
    snap = a.snapshot()
    a.advance(200.0)
    ref, bars = a.readState(), a.readBarriers()
    b = _rich_solver(lts, True)
    b.setRainField(np.ones(a.nx * a.ny))   # a different field: the snapshot's must come back
    b.restore(snap)
    b.rainRate = 120 * rc.MMHR
    b.advance(200.0)
    s = b.readState()
    for k in ("h", "hu", "hv", "hmax"):
        np.testing.assert_array_equal(s[k], ref[k])
    for k in ("volume", "rain", "infil", "outflow", "giStored", "expected"):
        assert s["mass"][k] == ref["mass"][k], k
    assert ref["mass"]["giStored"] > 0 and ref["mass"]["infil"] > 0
    np.testing.assert_equal(b.readBarriers(), bars)
    # a solver without the late rain field is another model
    import floodsim_amd as fsa
    c = _rich_solver(lts, False)
    with pytest.raises(fsa.FloodSimError):
        c.restore(snap)
    for o in (a, b, c):
        o.destroy()


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

"""FP64 CPU oracle for the shallow-water step -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package, and only as the checker: the product (floodsim_amd) never uses it.

`CPUSolver` wraps oracle.c (a C restatement of the reference's src/solver-cpu.js,
class CPUSolver, solver-cpu.js:29-576) through ctypes and keeps the reference's
method names: setDepth, setVelocity, run, step, field, massBalance.

Pinning (each function is checked by a test in tests/test_oracle.py):
  * MUSCL/HLLC/SSP-RK2/CFL: Stoker and Ritter dam breaks against the recorded L1
    errors, step counts and nx=400 profiles (tests/results-cpu.json) and against
    the closed form (src/analytic.js:5-28).
  * Audusse well-balancing: lake at rest (max |u|, |eta - eta0|, exactly 88 dry cells).
  * Wet/dry shoreline: Thacker L1 at T/4..T, step counts, centreline profile.
  * Rain, infiltration, open edge, walls, theta limiter: rain-on-grid books and
    step count; urban cloudburst and valley dam-break books and step counts.
  * Local time stepping: LTS tests of tests/run-lts-tests.mjs (mass, L1 within
    10 % of global, still water, >= 2.5x fewer updates) and the GPU LTS figures of
    results-gpu.json (cycles, L1).
  * Manning friction: sheet-flow kinematic-wave closed form
    (tests/run-sheetflow-tests.mjs:10-21 gives the ratio for the default scheme).
  * F26 options (implicitFriction + stepFlux): the same sheet flow within 12 % of the
    kinematic depth at 40, 80 and 100 m on two planes, mass balance under 1e-12
    (tests/run-sheetflow-tests.mjs:22-27); the implicit factor against its defining
    quadratic k = 1 + a |u / k|.
  * Outflow BC: never imports water on an inward-tilting slope
    (tests/run-phase1-tests.mjs:112-127).
  * Inlet exchange: exact books (exchanged volume closes the mass balance) and
    the rule of exec-cpu.mjs:46-55 checked cell by cell on a tiny case.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

BC = {"wall": 0, "open": 1, "outflow": 2}


class _Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("dx", ctypes.c_double), ("g", ctypes.c_double), ("cfl", ctypes.c_double),
        ("hdry", ctypes.c_double), ("hfront", ctypes.c_double), ("dtDry", ctypes.c_double),
        ("secondOrder", ctypes.c_int), ("limiter", ctypes.c_int),
        ("bc", ctypes.c_int * 4), ("lts", ctypes.c_int),
        ("implicitFriction", ctypes.c_int), ("stepFlux", ctypes.c_int),
    ]


def build():
    """Compile liboracle.so with the Makefile beside this file."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        D, I, P = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        L.or_create.restype = P
        L.or_create.argtypes = [ctypes.POINTER(_Params), dp, dp, D, ctypes.POINTER(ctypes.c_uint8)]
        L.or_destroy.argtypes = [P]
        for name in ("or_set_infil", "or_set_rain_field", "or_set_depth"):
            getattr(L, name).argtypes = [P, dp]
        L.or_set_velocity.argtypes = [P, dp, dp]
        L.or_set_state.argtypes = [P, dp, dp, dp]
        L.or_set_rain_rate.argtypes = [P, D]
        L.or_set_inlets.argtypes = [P, ctypes.POINTER(ctypes.c_int), I]
        L.or_set_stage_cells.argtypes = [P, ctypes.POINTER(ctypes.c_int), I]
        L.or_set_stage.argtypes = [P, D]
        ip = ctypes.POINTER(ctypes.c_int)
        L.or_set_crests.argtypes = [P, ip, ip, ip, dp, dp, dp, I]
        L.or_set_crest_levels.argtypes = [P, dp]
        L.or_read_barriers.argtypes = [P, dp]
        L.or_set_ga.argtypes = [P, dp, dp, dp]
        L.or_set_gi.argtypes = [P, dp, dp, dp]
        L.or_set_inlet_rates.argtypes = [P, dp]
        L.or_read_inlets.argtypes = [P, dp, dp]
        L.or_step.restype = D
        L.or_step.argtypes = [P, D]
        L.or_run.argtypes = [P, D]
        L.or_run_steps.restype = ctypes.c_long
        L.or_run_steps.argtypes = [P, D, ctypes.c_long]
        L.or_compute_dt.restype = D
        L.or_compute_dt.argtypes = [P]
        L.or_get_field.argtypes = [P, I, dp]
        L.or_get_hazard.argtypes = [P, I, dp]
        L.or_hazard_cell.argtypes = [P, I, D, D, D, D]
        L.or_hz_arr.argtypes = [ctypes.c_double, ctypes.c_double]
        L.or_hz_arr.restype = I
        L.or_hz_hr.argtypes = [ctypes.c_double]
        L.or_hz_hr.restype = I
        L.or_hz_force.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.or_hz_force.restype = ctypes.c_double
        L.or_get_levels.restype = I
        L.or_get_levels.argtypes = [P, ctypes.POINTER(ctypes.c_int)]
        L.or_books.argtypes = [P, dp]
        L.or_reset_books.argtypes = [P, D]
        L.or_volume.restype = D
        L.or_volume.argtypes = [P]
        L.or_set_threads.argtypes = [I]
        _LIB = L
    return _LIB


def set_threads(n):
    lib().or_set_threads(int(n))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


_FIELDS = {"h": 0, "hu": 1, "hv": 2, "maxDepth": 3, "hmax": 3, "z": 4, "aH": 5, "gaF": 6, "giStore": 7}


class CPUSolver:
    """The reference CPUSolver's options and methods (solver-cpu.js:52-132) over oracle.c.

    Options: nx, ny, dx, z (nx*ny), manning (number or array), mask (uint8, 0 = wall),
    bc {west,east,south,north}, secondOrder, limiter, cfl, hdry, hfront, dtDry, lts,
    infil (number or array, m/s), rainField (array), inlets (interior indices j*nx+i),
    implicitFriction, stepFlux (F26; stepFlux has no effect with lts, as in the reference),
    stageCells (interior edge cells whose outward ghosts see the level set by setStage),
    crest (flood-barrier faces [{i, j, axis, z, C, ground}], solver-cpu.js:43, 98-105),
    ga {psiD, fmax, perv} (Green-Ampt soil) and gi {rate, cap, drain} (GI stores), per cell
    (solver-cpu.js:94-96).
    """

    def __init__(self, nx, ny, dx, z, manning=0.0, mask=None, bc=None, secondOrder=True,
                 limiter="minmod", cfl=None, hdry=1e-6, hfront=1e-3, dtDry=1.0, lts=0,
                 g=9.81, infil=None, rainField=None, inlets=None, implicitFriction=False, stepFlux=False,
                 stageCells=None, crest=None, ga=None, gi=None, **_ignored):
        L = lib()
        self.nx, self.ny, self.dx = int(nx), int(ny), float(dx)
        b = {"west": "wall", "east": "wall", "south": "wall", "north": "wall"}
        b.update(bc or {})
        self.bc = b
        p = _Params()
        p.nx, p.ny, p.dx, p.g = self.nx, self.ny, self.dx, g
        p.cfl = cfl if cfl is not None else (0.48 if secondOrder else 0.9)
        p.hdry, p.hfront, p.dtDry = hdry, hfront, dtDry
        p.secondOrder = 1 if secondOrder else 0
        p.limiter = 1 if limiter == "vanleer" else 0
        for k, side in enumerate(("west", "east", "south", "north")):
            p.bc[k] = BC[b[side]]
        if lts is True:
            lts = 3
        p.lts = int(lts or 0)
        p.implicitFriction = 1 if implicitFriction else 0   # F26 (solver-cpu.js:62-63)
        p.stepFlux = 1 if stepFlux else 0
        self.lts = p.lts
        self.hdry, self.cfl = hdry, p.cfl
        n = self.nx * self.ny
        self._z = np.ascontiguousarray(np.asarray(z, dtype=np.float64).reshape(n))
        if isinstance(manning, (int, float)):
            man_arr, man_s = None, float(manning)
        else:
            man_arr, man_s = np.ascontiguousarray(np.asarray(manning, dtype=np.float64).reshape(n)), 0.0
        self._mask = None if mask is None else np.ascontiguousarray(np.asarray(mask, dtype=np.uint8).reshape(n))
        self._h = L.or_create(ctypes.byref(p), _dp(self._z), None if man_arr is None else _dp(man_arr), man_s,
                              None if self._mask is None else self._mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
        if infil is not None:
            self.setInfil(infil)
        if rainField is not None:
            rf = np.ascontiguousarray(np.asarray(rainField, dtype=np.float64).reshape(n))
            L.or_set_rain_field(self._h, _dp(rf))
        self.nInlets = 0
        if inlets is not None and len(inlets):
            ink = np.ascontiguousarray(np.asarray(inlets, dtype=np.int32))
            self.nInlets = len(ink)
            L.or_set_inlets(self._h, ink.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(ink))
        self.nBar = 0
        if crest:
            self.nBar = len(crest)
            col = lambda k, d, t: np.ascontiguousarray([f.get(k, d) if f.get(k, d) is not None else d for f in crest],
                                                       dtype=t)
            self._bi, self._bj, self._ba = col("i", 0, np.int32), col("j", 0, np.int32), col("axis", 0, np.int32)
            bz, bc, bg = col("z", -np.inf, np.float64), col("C", 1.6, np.float64), col("ground", -np.inf, np.float64)
            ipt = ctypes.POINTER(ctypes.c_int)
            L.or_set_crests(self._h, self._bi.ctypes.data_as(ipt), self._bj.ctypes.data_as(ipt),
                            self._ba.ctypes.data_as(ipt), _dp(bz), _dp(bc), _dp(bg), self.nBar)
        if ga is not None:
            a = [np.ascontiguousarray(np.asarray(ga[k], dtype=np.float64).reshape(n)) for k in ("psiD", "fmax", "perv")]
            L.or_set_ga(self._h, *[_dp(x) for x in a])
        if gi is not None:
            a = [np.ascontiguousarray(np.asarray(gi[k], dtype=np.float64).reshape(n)) for k in ("rate", "cap", "drain")]
            L.or_set_gi(self._h, *[_dp(x) for x in a])
        if stageCells is not None and len(stageCells):
            sk = np.ascontiguousarray(np.asarray(stageCells, dtype=np.int32))
            L.or_set_stage_cells(self._h, sk.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(sk))
        self._rain = 0.0

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_destroy(self._h)
            self._h = None

    # --- inputs
    def _centres(self):
        x = (np.arange(self.nx) + 0.5) * self.dx
        y = (np.arange(self.ny) + 0.5) * self.dx
        return np.meshgrid(x, y)

    def setInfil(self, infil):
        n = self.nx * self.ny
        a = np.full(n, float(infil)) if np.isscalar(infil) else np.asarray(infil, dtype=np.float64).reshape(n)
        a = np.ascontiguousarray(a)
        lib().or_set_infil(self._h, _dp(a))

    def setDepth(self, fn_or_array):
        """fn(x, y, z) at cell centres (vectorised over numpy arrays), or an nx*ny array."""
        if callable(fn_or_array):
            X, Y = self._centres()
            h = fn_or_array(X, Y, self._z.reshape(self.ny, self.nx))
            h = np.broadcast_to(np.asarray(h, dtype=np.float64), (self.ny, self.nx))
        else:
            h = np.asarray(fn_or_array, dtype=np.float64)
        h = np.ascontiguousarray(h.reshape(self.nx * self.ny))
        lib().or_set_depth(self._h, _dp(h))

    def setRainField(self, field):
        """A new per-cell rain multiplier (the reference CPU's rainField property, solver-cpu.js:71, 470)."""
        rf = np.ascontiguousarray(np.asarray(field, dtype=np.float64).reshape(self.nx * self.ny))
        lib().or_set_rain_field(self._h, _dp(rf))

    def setVelocity(self, fn):
        X, Y = self._centres()
        u, v = fn(X, Y)
        u = np.ascontiguousarray(np.broadcast_to(np.asarray(u, dtype=np.float64), X.shape).reshape(-1))
        v = np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), X.shape).reshape(-1))
        lib().or_set_velocity(self._h, _dp(u), _dp(v))

    def setState(self, h, hu, hv, t=None):
        a = [np.ascontiguousarray(np.asarray(q, dtype=np.float64).reshape(-1)) for q in (h, hu, hv)]
        lib().or_set_state(self._h, _dp(a[0]), _dp(a[1]), _dp(a[2]))
        if t is not None:
            lib().or_reset_books(self._h, float(t))

    @property
    def rainRate(self):
        return self._rain

    @rainRate.setter
    def rainRate(self, r):
        self._rain = float(r)
        lib().or_set_rain_rate(self._h, self._rain)

    def setCrestLevels(self, zs):
        """New crest per barrier face, in the order of `crest` (-inf = not there; solver-cpu.js:136-138)."""
        a = np.ascontiguousarray(np.asarray(zs, dtype=np.float64).reshape(self.nBar))
        lib().or_set_crest_levels(self._h, _dp(a))

    def readBarriers(self):
        """Barrier books per face (solver-cpu.js:139): plus, minus (m^3), peak (m^2/s), level (m), load (kN/m),
        tPeak, tLevel, tLoad, tOver (s)."""
        a = np.zeros(9 * max(self.nBar, 1))
        if self.nBar:
            lib().or_read_barriers(self._h, _dp(a))
        a = a[:9 * self.nBar].reshape(self.nBar, 9)
        return {k: a[:, q].copy() for q, k in enumerate(("plus", "minus", "peak", "level", "load", "tPeak", "tLevel",
                                                          "tLoad", "tOver"))}

    def setStage(self, z):
        """The water level the stage cells' ghosts hold (m, absolute; solver-cpu.js:204)."""
        lib().or_set_stage(self._h, float(z))

    def setInletRates(self, q):
        a = np.ascontiguousarray(np.asarray(q, dtype=np.float64))
        lib().or_set_inlet_rates(self._h, _dp(a))

    def readInlets(self):
        vol = np.zeros(self.nInlets)
        h = np.zeros(self.nInlets)
        lib().or_read_inlets(self._h, _dp(vol), _dp(h))
        return {"vol": vol, "h": h}

    # --- stepping
    def step(self, dtMax=float("inf")):
        return lib().or_step(self._h, float(dtMax))

    def run(self, tEnd):
        lib().or_run(self._h, float(tEnd))

    def run_steps(self, tEnd, k):
        return lib().or_run_steps(self._h, float(tEnd), int(k))

    def computeDt(self):
        return lib().or_compute_dt(self._h)

    # --- outputs
    def field(self, name):
        out = np.empty(self.nx * self.ny)
        lib().or_get_field(self._h, _FIELDS[name], _dp(out))
        return out

    def readHazard(self):
        """Hazard products per interior cell (solver-cpu.js:140 readHazard; hazard.js:75-96)."""
        n = self.nx * self.ny
        names = ("dv", "force", "duration", "arrival", "vmax", "tWet", "arr", "defra", "defraUrban")
        out = {}
        for k, name in enumerate(names):
            a = np.zeros(n)
            lib().or_get_hazard(self._h, k, _dp(a))
            out[name] = a if k < 6 else a.astype(np.uint8)
        return out

    def levels(self):
        tx, ty = -(-self.nx // 8), -(-self.ny // 8)
        out = np.zeros(tx * ty, dtype=np.int32)
        lib().or_get_levels(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return out.reshape(ty, tx)

    def _books(self):
        o = np.zeros(13)
        lib().or_books(self._h, _dp(o))
        return o

    @property
    def t(self):
        return self._books()[0]

    @property
    def steps(self):
        return int(self._books()[1])

    @property
    def cellSteps(self):
        return int(self._books()[11])

    @property
    def records(self):
        return int(self._books()[12])

    @property
    def vol(self):
        o = self._books()
        return {"initial": o[3], "rain": o[4], "infil": o[5], "outflow": o[6], "clipped": o[7], "exchanged": o[8]}

    def volume(self):
        return lib().or_volume(self._h)

    def massBalance(self):
        o = self._books()
        return {"volume": o[2], "expected": o[9], "error": o[2] - o[9], "rel": o[10], "clipped": o[7]}

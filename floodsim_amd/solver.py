"""ctypes binding of libfloodsim_hip.so (include/floodsim_hip.h).

`GPUSolver` keeps the reference's solver-object interface (src/solver-gpu.js:13-693): built from
the model layers (constructor options of solver-gpu.js:35-47), advanced in time (advance, frame),
depths and velocities read back (readState), plus rainRate, setRainField, setInletRates,
readInlets, tileStats and destroy. The JavaScript twin for Node is floodsim_amd/js/solver-hip.js.

`Pool` is the batched/served form: many instances advanced by one launch sequence.

There is no fallback: if the HIP library is missing or no device is present, construction raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("FS_LIB_PATH") or os.path.join(_HERE, "libfloodsim_hip.so")
_LIB = None

BC = {"wall": 0, "open": 1, "outflow": 2}
MMHR = 1 / 1000 / 3600


class FloodSimError(RuntimeError):
    pass


class fs_params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("dx", ctypes.c_double), ("g", ctypes.c_double), ("cfl", ctypes.c_double),
        ("hdry", ctypes.c_double), ("hfront", ctypes.c_double), ("dtDry", ctypes.c_double),
        ("secondOrder", ctypes.c_int32), ("limiter", ctypes.c_int32),
        ("bc", ctypes.c_int32 * 4), ("lts", ctypes.c_int32), ("rows", ctypes.c_int32),
        ("untiled", ctypes.c_int32), ("hazard", ctypes.c_int32),
        ("implicitFriction", ctypes.c_int32), ("stepFlux", ctypes.c_int32),
    ]


class fs_clock(ctypes.Structure):
    _fields_ = [("t", ctypes.c_double), ("lastDt", ctypes.c_double), ("dtAvg", ctypes.c_double),
                ("steps", ctypes.c_int64), ("cellUpdates", ctypes.c_int64)]


class fs_tiles(ctypes.Structure):
    _fields_ = [("tiled", ctypes.c_int32), ("pad", ctypes.c_int32)] + \
        [(k, ctypes.c_int64) for k in ("tiles", "update", "flux", "cellsActive", "cellsTotal")]


class fs_mass(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in
                ("volume", "rain", "infil", "outflow", "exchanged", "clipped", "expected", "rel", "relExClip", "initial",
                 "giStored")]


EXPORTS = [
    "fs_create", "fs_destroy", "fs_last_error", "fs_set_state", "fs_set_rain", "fs_set_rain_field",
    "fs_set_inlet_rates", "fs_read_inlets", "fs_read_inlet_approach", "fs_advance", "fs_frame", "fs_read_state", "fs_clock_get",
    "fs_tile_stats", "fs_pool_create", "fs_pool_destroy", "fs_pool_last_error", "fs_pool_bytes_needed",
    "fs_pool_add", "fs_pool_remove", "fs_pool_set_state", "fs_pool_set_rain", "fs_pool_set_inlet_rates",
    "fs_pool_advance", "fs_pool_step", "fs_pool_read_state", "fs_pool_clock", "fs_pool_bytes_used",
    "fs_pool_kernel_ms", "fs_pool_set_timing", "fs_solver_pool", "fs_version", "fs_copy_probe",
    "fs_pool_dominant_kernel", "fs_lts_levels", "fs_pool_set_rows", "fs_snapshot_bytes", "fs_snapshot",
    "fs_restore", "fs_pool_snapshot_bytes", "fs_pool_snapshot", "fs_pool_restore", "fs_read_hazard",
    "fs_read_infiltration",
    "fs_set_stage_cells", "fs_set_stage", "fs_pool_set_stage_cells", "fs_pool_set_stage",
    "fs_set_crests", "fs_set_crest_levels", "fs_read_barriers", "fs_pool_set_crests", "fs_pool_set_crest_levels",
    "fs_pool_read_barriers", "fs_set_ga", "fs_set_gi", "fs_pool_set_ga", "fs_pool_set_gi", "fs_pool_debug",
    "fs_advance_noalloc", "fs_snapshot_layout",
]
SNAP_PARTS = ("record", "planes", "barriers", "soil", "rainField")   # fs_snapshot_layout's parts, in order
FS_DEBUG_NO_STEPS = 1


def lib(path=None):
    """Load the HIP library (raises FloodSimError when it is not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FloodSimError(f"{p} is not built: run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    L = ctypes.CDLL(p)
    P, I32, D, SZ = ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, ctypes.c_size_t
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int32)
    u8p = ctypes.POINTER(ctypes.c_uint8)
    pp = ctypes.POINTER(fs_params)
    L.fs_version.restype = ctypes.c_char_p
    L.fs_last_error.restype = ctypes.c_char_p
    L.fs_last_error.argtypes = [P]
    L.fs_pool_last_error.restype = ctypes.c_char_p
    L.fs_pool_last_error.argtypes = [P]
    L.fs_create.argtypes = [pp, dp, dp, D, u8p, dp, dp, ip, I32, ctypes.POINTER(P)]
    L.fs_destroy.argtypes = [P]
    L.fs_destroy.restype = None
    L.fs_set_state.argtypes = [P, dp, dp, dp, I32]
    L.fs_set_rain.argtypes = [P, D]
    L.fs_set_rain_field.argtypes = [P, dp]
    L.fs_set_inlet_rates.argtypes = [P, dp, I32]
    L.fs_read_inlets.argtypes = [P, dp, dp, I32]
    L.fs_read_inlet_approach.argtypes = [P, ctypes.POINTER(ctypes.c_float), I32]
    L.fs_advance.argtypes = [P, D, I32, ctypes.POINTER(fs_clock)]
    L.fs_advance_noalloc.argtypes = [P, D, I32, ctypes.POINTER(fs_clock)]
    L.fs_frame.argtypes = [P, D, I32, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(fs_clock)]
    L.fs_read_state.argtypes = [P, dp, dp, dp, dp, ctypes.POINTER(fs_mass)]
    L.fs_clock_get.argtypes = [P, ctypes.POINTER(fs_clock)]
    L.fs_tile_stats.argtypes = [P, ctypes.POINTER(fs_tiles)]
    L.fs_pool_create.argtypes = [SZ, ctypes.POINTER(P)]
    L.fs_pool_destroy.argtypes = [P]
    L.fs_pool_destroy.restype = None
    L.fs_pool_bytes_needed.restype = SZ
    L.fs_pool_bytes_needed.argtypes = [pp, I32, I32]
    L.fs_pool_bytes_used.restype = SZ
    L.fs_pool_bytes_used.argtypes = [P]
    L.fs_pool_add.argtypes = [P, pp, dp, dp, D, u8p, dp, dp, ip, I32, ip]
    L.fs_pool_remove.argtypes = [P, I32]
    L.fs_pool_set_rows.argtypes = [P, I32, I32]
    L.fs_pool_set_state.argtypes = [P, I32, dp, dp, dp, I32]
    L.fs_pool_set_rain.argtypes = [P, I32, D]
    L.fs_pool_set_inlet_rates.argtypes = [P, I32, dp, I32]
    L.fs_pool_set_stage_cells.argtypes = [P, I32, ip, I32]
    L.fs_pool_set_stage.argtypes = [P, I32, D]
    L.fs_set_stage_cells.argtypes = [P, ip, I32]
    L.fs_set_stage.argtypes = [P, D]
    L.fs_set_crests.argtypes = [P, ip, ip, ip, dp, dp, dp, I32]
    L.fs_set_crest_levels.argtypes = [P, dp]
    L.fs_read_barriers.argtypes = [P, dp]
    L.fs_pool_set_crests.argtypes = [P, I32, ip, ip, ip, dp, dp, dp, I32]
    L.fs_pool_set_crest_levels.argtypes = [P, I32, dp]
    L.fs_pool_read_barriers.argtypes = [P, I32, dp]
    L.fs_set_ga.argtypes = [P, dp, dp, dp]
    L.fs_set_gi.argtypes = [P, dp, dp, dp]
    L.fs_pool_set_ga.argtypes = [P, I32, dp, dp, dp]
    L.fs_pool_set_gi.argtypes = [P, I32, dp, dp, dp]
    L.fs_pool_advance.argtypes = [P, ip, dp, I32, I32, ctypes.POINTER(fs_clock)]
    L.fs_pool_step.argtypes = [P, ip, dp, I32, I32]
    L.fs_pool_read_state.argtypes = [P, I32, dp, dp, dp, dp, ctypes.POINTER(fs_mass)]
    L.fs_pool_clock.argtypes = [P, I32, ctypes.POINTER(fs_clock)]
    L.fs_pool_kernel_ms.argtypes = [P, dp, dp, ctypes.POINTER(ctypes.c_int64)]
    L.fs_pool_set_timing.argtypes = [P, I32]
    L.fs_pool_debug.argtypes = [P, I32, I32]
    L.fs_copy_probe.argtypes = [SZ, I32, dp, dp]
    L.fs_lts_levels.argtypes = [P, ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
    L.fs_pool_dominant_kernel.argtypes = [P, dp, dp, ctypes.POINTER(ctypes.c_int64)]
    for f in (L.fs_snapshot_bytes, L.fs_pool_snapshot_bytes):
        f.restype = SZ
    L.fs_snapshot_bytes.argtypes = [P]
    L.fs_pool_snapshot_bytes.argtypes = [P, I32]
    L.fs_snapshot.argtypes = [P, P, SZ]
    L.fs_restore.argtypes = [P, P, SZ]
    L.fs_pool_snapshot.argtypes = [P, I32, P, SZ]
    L.fs_pool_restore.argtypes = [P, I32, P, SZ]
    L.fs_snapshot_layout.argtypes = [P, ctypes.POINTER(ctypes.c_uint64)]
    fp = ctypes.POINTER(ctypes.c_float)
    L.fs_read_hazard.argtypes = [P, fp, fp, fp, fp, fp, fp, u8p, u8p, u8p]
    L.fs_read_infiltration.argtypes = [P, fp]
    L.fs_solver_pool.restype = P
    L.fs_solver_pool.argtypes = [P]
    _LIB = L
    return L


def _dp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f64(a, n):
    return None if a is None else np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (n,)) if np.isscalar(a) else np.asarray(a, dtype=np.float64).reshape(n))


def make_params(nx, ny, dx, bc=None, secondOrder=True, limiter="minmod", cfl=None, hdry=1e-5, hfront=1e-3,
                dtDry=1.0, g=9.81, lts=0, rows=0, tiles=True, hazard=False, implicitFriction=False,
                stepFlux=False, **_):
    """fs_params with the defaults of GPUSolver (solver-gpu.js:39-47)."""
    p = fs_params()
    p.nx, p.ny, p.dx, p.g = int(nx), int(ny), float(dx), float(g)
    p.cfl = float(cfl) if cfl is not None else (0.48 if secondOrder else 0.9)
    p.hdry, p.hfront, p.dtDry = float(hdry), float(hfront), float(dtDry)
    p.secondOrder = 1 if secondOrder else 0
    p.limiter = 1 if limiter == "vanleer" else 0
    b = {"west": "wall", "east": "wall", "south": "wall", "north": "wall"}
    b.update(bc or {})
    for k, side in enumerate(("west", "east", "south", "north")):
        p.bc[k] = BC[b[side]]
    if lts is True:
        lts = 3
    p.lts = int(lts or 0)
    p.rows = int(rows or 0)
    p.untiled = 0 if tiles else 1   # tiles=False: every strip every step (the reference's o.tiles)
    p.hazard = 1 if hazard else 0   # hazard products every step (readHazard)
    p.implicitFriction = 1 if implicitFriction else 0   # F26 (solver-gpu.js:43)
    p.stepFlux = 1 if stepFlux else 0
    return p


class _Layers:
    """Host-side model layers in the C-ABI's layout (interior nx*ny, row-major, FP64)."""

    def __init__(self, nx, ny, z, manning=0.0, mask=None, infil=None, rainField=None, inlets=None):
        n = nx * ny
        self.z = _f64(z, n)
        if np.isscalar(manning) or manning is None:
            self.man, self.man_s = None, float(manning or 0.0)
        else:
            self.man, self.man_s = _f64(manning, n), 0.0
        self.mask = None if mask is None else np.ascontiguousarray(np.asarray(mask, dtype=np.uint8).reshape(n))
        self.infil = None if infil is None else _f64(infil, n)
        self.rf = None if rainField is None else _f64(rainField, n)
        self.inlets = np.ascontiguousarray(np.asarray(inlets if inlets is not None else [], dtype=np.int32))

    def args(self):
        return (_dp(self.z), _dp(self.man), self.man_s,
                None if self.mask is None else self.mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                _dp(self.infil), _dp(self.rf),
                self.inlets.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)) if len(self.inlets) else None,
                len(self.inlets))


def _crest_args(crest):
    """fs_set_crests arguments from the reference's crest list [{i, j, axis, z, C, ground}]."""
    col = lambda k, d, t: np.ascontiguousarray([d if f.get(k) is None else f.get(k) for f in crest], dtype=t)
    a = [col("i", 0, np.int32), col("j", 0, np.int32), col("axis", 0, np.int32),
         col("z", -np.inf, np.float64), col("C", 1.6, np.float64), col("ground", -np.inf, np.float64)]
    _crest_args.keep = a
    ip = ctypes.POINTER(ctypes.c_int32)
    return [a[0].ctypes.data_as(ip), a[1].ctypes.data_as(ip), a[2].ctypes.data_as(ip), _dp(a[3]), _dp(a[4]),
            _dp(a[5]), len(crest)]


def _barrier_books(read, n):
    out = np.zeros(9 * max(n, 1))
    if n:
        read(_dp(out))
    out = out[:9 * n].reshape(n, 9)
    return {k: out[:, q].copy() for q, k in enumerate(("plus", "minus", "peak", "level", "load", "tPeak", "tLevel",
                                                        "tLoad", "tOver"))}


class GPUSolver:
    """The reference GPUSolver interface over the HIP library (one instance)."""

    def __init__(self, nx, ny, dx, z, manning=0.0, mask=None, infil=None, rainField=None, inlets=None,
                 stageCells=None, crest=None, ga=None, gi=None, **opts):
        L = lib()
        self.nx, self.ny, self.dx = int(nx), int(ny), float(dx)
        self.p = make_params(nx, ny, dx, **opts)
        self.hdry = self.p.hdry
        self.layers = _Layers(self.nx, self.ny, z, manning, mask, infil, rainField, inlets)
        if np.isscalar(infil) and infil is not None:
            self.layers.infil = _f64(float(infil), self.nx * self.ny)
        h = ctypes.c_void_p()
        rc = L.fs_create(ctypes.byref(self.p), *self.layers.args(), ctypes.byref(h))
        if rc != 0:
            raise FloodSimError(f"fs_create failed ({rc}): {L.fs_last_error(None).decode()}")
        self._h = h
        if stageCells is not None and len(stageCells):
            sk = np.ascontiguousarray(np.asarray(stageCells, dtype=np.int32))
            self._chk(L.fs_set_stage_cells(h, sk.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(sk)))
        self.nBar = 0
        if crest:
            self.nBar = len(crest)
            self._chk(L.fs_set_crests(h, *_crest_args(crest)))
        n = self.nx * self.ny
        if ga is not None:
            self._chk(L.fs_set_ga(h, *[_dp(_f64(ga[k], n)) for k in ("psiD", "fmax", "perv")]))
        if gi is not None:
            self._chk(L.fs_set_gi(h, *[_dp(_f64(gi[k], n)) for k in ("rate", "cap", "drain")]))
        self.nInlets = len(self.layers.inlets)
        self._rain = 0.0
        self._Uh = np.zeros(self.nx * self.ny)
        self._hu = np.zeros(self.nx * self.ny)
        self._hv = np.zeros(self.nx * self.ny)
        self.zref = float(self.layers.z.min())

    def _chk(self, rc):
        if rc != 0:
            raise FloodSimError(f"floodsim_hip error {rc}: {lib().fs_last_error(self._h).decode()}")

    def destroy(self):
        if getattr(self, "_h", None):
            lib().fs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    # --- state (setDepth / setVelocity, solver-gpu.js:301-318)
    def _centres(self):
        x = (np.arange(self.nx) + 0.5) * self.dx
        y = (np.arange(self.ny) + 0.5) * self.dx
        return np.meshgrid(x, y)

    def setDepth(self, fn_or_array):
        if callable(fn_or_array):
            X, Y = self._centres()
            h = fn_or_array(X, Y, self.layers.z.reshape(self.ny, self.nx))
        else:
            h = fn_or_array
        h = np.asarray(h, dtype=np.float64)
        h = h.reshape(-1) if h.size == self.nx * self.ny else np.broadcast_to(h, (self.ny, self.nx)).reshape(-1)
        if self.layers.mask is not None:
            h = np.where(self.layers.mask != 0, np.maximum(0, h), 0.0)
        else:
            h = np.maximum(0, h)
        self._Uh = np.ascontiguousarray(h.astype(np.float32).astype(np.float64))
        self._hu[:] = 0
        self._hv[:] = 0
        self._chk(lib().fs_set_state(self._h, _dp(self._Uh), _dp(self._hu), _dp(self._hv), 1))

    def setVelocity(self, fn):
        X, Y = self._centres()
        u, v = fn(X, Y)
        u = np.broadcast_to(np.asarray(u, dtype=np.float64), X.shape).reshape(-1)
        v = np.broadcast_to(np.asarray(v, dtype=np.float64), X.shape).reshape(-1)
        # hu = h * u with the stored FP32 depth, rounded once to FP32 (solver-gpu.js:315)
        self._hu = np.ascontiguousarray(self._Uh * u)
        self._hv = np.ascontiguousarray(self._Uh * v)
        self._chk(lib().fs_set_state(self._h, _dp(self._Uh), _dp(self._hu), _dp(self._hv), 1))

    def setState(self, h, hu, hv, bookInitial=True):
        a = [np.ascontiguousarray(np.asarray(q, dtype=np.float64).reshape(-1)) for q in (h, hu, hv)]
        self._chk(lib().fs_set_state(self._h, _dp(a[0]), _dp(a[1]), _dp(a[2]), 1 if bookInitial else 0))

    @property
    def rainRate(self):
        return self._rain

    @rainRate.setter
    def rainRate(self, r):
        self._rain = float(r)
        self._chk(lib().fs_set_rain(self._h, self._rain))

    def setRainField(self, field):
        f = _f64(field, self.nx * self.ny)
        self._chk(lib().fs_set_rain_field(self._h, _dp(f)))

    def setStage(self, z):
        """Water level held by the stage cells' ghosts (m, absolute; solver-gpu.js:554)."""
        self._chk(lib().fs_set_stage(self._h, float(z)))

    def setCrestLevels(self, zs):
        """New crest per barrier face, in the order of `crest` (-inf = not there; solver-gpu.js:486-488)."""
        a = _f64(zs, self.nBar)
        self._chk(lib().fs_set_crest_levels(self._h, _dp(a)))

    def readBarriers(self):
        """Barrier books per face (solver-gpu.js:490-504): plus, minus, peak, level, load, tPeak, tLevel,
        tLoad, tOver."""
        return _barrier_books(lambda out: self._chk(lib().fs_read_barriers(self._h, out)), self.nBar)

    def readInfiltration(self):
        """Depth infiltrated so far per interior cell under Green-Ampt (m per pervious area; 0 without it;
        solver-gpu.js:507-512)."""
        F = np.zeros(self.nx * self.ny, dtype=np.float32)
        self._chk(lib().fs_read_infiltration(self._h, F.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return F

    def setInletRates(self, q):
        a = np.ascontiguousarray(np.asarray(q, dtype=np.float64))
        self._chk(lib().fs_set_inlet_rates(self._h, _dp(a), len(a)))

    def readInlets(self):
        vol = np.zeros(self.nInlets)
        h = np.zeros(self.nInlets)
        self._chk(lib().fs_read_inlets(self._h, _dp(vol), _dp(h), self.nInlets))
        app = np.zeros(4 * self.nInlets, dtype=np.float32)
        self._chk(lib().fs_read_inlet_approach(self._h, app.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                               self.nInlets))
        return {"vol": vol, "h": h, "app": app}

    # --- stepping (advance / frame, solver-gpu.js:589-622)
    def advance(self, tTarget, K=64):
        c = fs_clock()
        self._chk(lib().fs_advance(self._h, float(tTarget), int(K), ctypes.byref(c)))
        return c

    def frame(self, tEnd, K, readOut=True):
        out = np.zeros(self.nx * self.ny * 2, dtype=np.float32) if readOut else None
        c = fs_clock()
        self._chk(lib().fs_frame(self._h, float(tEnd), int(K),
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if readOut else None, ctypes.byref(c)))
        return out

    def clock(self):
        c = fs_clock()
        self._chk(lib().fs_clock_get(self._h, ctypes.byref(c)))
        return c

    @property
    def t(self):
        return self.clock().t

    @property
    def steps(self):
        return self.clock().steps

    # --- readback (readState, solver-gpu.js:651-690)
    def readState(self):
        n = self.nx * self.ny
        h, hu, hv, hm = (np.zeros(n) for _ in range(4))
        m = fs_mass()
        self._chk(lib().fs_read_state(self._h, _dp(h), _dp(hu), _dp(hv), _dp(hm), ctypes.byref(m)))
        mass = {k: getattr(m, k) for k, _ in fs_mass._fields_}
        return {"h": h, "hu": hu, "hv": hv, "hmax": hm, "mass": mass}

    def readHazard(self):
        """Hazard products per interior cell (readHazard, solver-gpu.js:515-529); needs hazard=True."""
        n = self.nx * self.ny
        f = {k: np.zeros(n, dtype=np.float32) for k in ("dv", "force", "duration", "arrival", "vmax", "tWet")}
        c = {k: np.zeros(n, dtype=np.uint8) for k in ("arr", "defra", "defraUrban")}
        fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        up = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        self._chk(lib().fs_read_hazard(self._h, fp(f["dv"]), fp(f["force"]), fp(f["duration"]), fp(f["arrival"]),
                                       fp(f["vmax"]), fp(f["tWet"]), up(c["arr"]), up(c["defra"]),
                                       up(c["defraUrban"])))
        f["cls"] = c
        return f

    # --- device-state snapshot / restore (checkpoint.js:94-113)
    def snapshot(self):
        """Everything the solver holds on the device plus its clock and books, as a uint8 array."""
        n = lib().fs_snapshot_bytes(self._h)
        buf = np.empty(n, dtype=np.uint8)
        self._chk(lib().fs_snapshot(self._h, buf.ctypes.data_as(ctypes.c_void_p), n))
        return buf

    def restore(self, buf):
        """Put a snapshot of this solver (or of one built from the same model) back."""
        b = np.ascontiguousarray(np.frombuffer(buf, dtype=np.uint8) if not isinstance(buf, np.ndarray) else buf)
        self._chk(lib().fs_restore(self._h, b.ctypes.data_as(ctypes.c_void_p), b.nbytes))

    def snapshot_layout(self):
        """Bytes of each named part of a snapshot, in order ({record, planes, barriers, soil, rainField}; a part
        the solver does not hold is 0)."""
        b = (ctypes.c_uint64 * len(SNAP_PARTS))()
        self._chk(lib().fs_snapshot_layout(self._h, b))
        return dict(zip(SNAP_PARTS, (int(x) for x in b)))

    def levels(self):
        """LTS tile levels of the last cycle, (ty, tx) uint8."""
        tx, ty = ctypes.c_int32(), ctypes.c_int32()
        self._chk(lib().fs_lts_levels(self._h, None, ctypes.byref(tx), ctypes.byref(ty)))
        out = np.zeros(tx.value * ty.value, dtype=np.uint8)
        self._chk(lib().fs_lts_levels(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), None, None))
        return out.reshape(ty.value, tx.value)

    def tileStats(self):
        """Active tiles of the last step (tileStats, solver-gpu.js:625-629): {tiled, tiles, update, flux} as
        tile counts (this build's skipping unit is the wavefront strip), plus the open cells the last step
        advanced (active) and the interior cell total (cells)."""
        t = fs_tiles()
        self._chk(lib().fs_tile_stats(self._h, ctypes.byref(t)))
        out = {"tiled": bool(t.tiled), "tiles": t.tiles, "active": t.cellsActive, "cells": t.cellsTotal}
        if t.tiled:
            out.update(update=t.update, flux=t.flux)
        return out

    def _debug_no_steps(self, on):
        """Test switch: every advance runs no steps (the stall path), until switched off."""
        self._chk(lib().fs_pool_debug(lib().fs_solver_pool(self._h), FS_DEBUG_NO_STEPS, 1 if on else 0))

    # --- timing of the stage kernels (HIP events on the solver's stream)
    def set_timing(self, on):
        self._chk(lib().fs_pool_set_timing(lib().fs_solver_pool(self._h), 1 if on else 0))

    def kernel_ms(self):
        a, b, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        self._chk(lib().fs_pool_kernel_ms(lib().fs_solver_pool(self._h), ctypes.byref(a), ctypes.byref(b), ctypes.byref(n)))
        return a.value, b.value, n.value

    def dominant_kernel(self):
        """(total ms, open cells advanced, launches) of the second-stage kernel since the last call."""
        a, b, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        self._chk(lib().fs_pool_dominant_kernel(lib().fs_solver_pool(self._h), ctypes.byref(a), ctypes.byref(b),
                                                ctypes.byref(n)))
        return a.value, b.value, n.value


def copy_probe(nbytes=2 << 30, iters=20):
    """Achievable HBM bandwidth (read + write GB/s) of a plain copy kernel, and its ms per copy."""
    g, ms = ctypes.c_double(), ctypes.c_double()
    rc = lib().fs_copy_probe(int(nbytes), int(iters), ctypes.byref(g), ctypes.byref(ms))
    if rc != 0:
        raise FloodSimError(f"fs_copy_probe failed ({rc}): {lib().fs_last_error(None).decode()}")
    return g.value, ms.value


class Pool:
    """Batched / served instances: one launch sequence advances many uneven instances."""

    def __init__(self, budget_bytes=0):
        L = lib()
        h = ctypes.c_void_p()
        rc = L.fs_pool_create(int(budget_bytes), ctypes.byref(h))
        if rc != 0:
            raise FloodSimError(f"fs_pool_create failed ({rc}): {L.fs_pool_last_error(None).decode()}")
        self._p = h
        self.meta = {}
        self.nbar = {}

    def _chk(self, rc):
        if rc != 0:
            raise FloodSimError(f"floodsim_hip pool error {rc}: {lib().fs_pool_last_error(self._p).decode()}")
        return rc

    def destroy(self):
        if getattr(self, "_p", None):
            lib().fs_pool_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    @staticmethod
    def bytes_needed(nx, ny, infil=False, rainField=False):
        p = make_params(nx, ny, 1.0)
        return lib().fs_pool_bytes_needed(ctypes.byref(p), 1 if infil else 0, 1 if rainField else 0)

    def bytes_used(self):
        return lib().fs_pool_bytes_used(self._p)

    def add(self, nx, ny, dx, z, manning=0.0, mask=None, infil=None, rainField=None, inlets=None, stageCells=None,
            crest=None, ga=None, gi=None, **opts):
        """Admit one instance; returns None when the HBM budget is spent."""
        p = make_params(nx, ny, dx, **opts)
        lay = _Layers(int(nx), int(ny), z, manning, mask, infil, rainField, inlets)
        slot = ctypes.c_int32()
        rc = lib().fs_pool_add(self._p, ctypes.byref(p), *lay.args(), ctypes.byref(slot))
        if rc == -3:
            return None
        self._chk(rc)
        self.meta[slot.value] = (int(nx), int(ny), float(dx), lay)
        if stageCells is not None and len(stageCells):
            self.set_stage_cells(slot.value, stageCells)
        if crest:
            self._chk(lib().fs_pool_set_crests(self._p, slot.value, *_crest_args(crest)))
            self.nbar[slot.value] = len(crest)
        n = int(nx) * int(ny)
        if ga is not None:
            self._chk(lib().fs_pool_set_ga(self._p, slot.value, *[_dp(_f64(ga[k], n)) for k in ("psiD", "fmax", "perv")]))
        if gi is not None:
            self._chk(lib().fs_pool_set_gi(self._p, slot.value, *[_dp(_f64(gi[k], n)) for k in ("rate", "cap", "drain")]))
        return slot.value

    def set_crest_levels(self, slot, zs):
        self._chk(lib().fs_pool_set_crest_levels(self._p, int(slot), _dp(_f64(zs, self.nbar[slot]))))

    def read_barriers(self, slot):
        return _barrier_books(lambda out: self._chk(lib().fs_pool_read_barriers(self._p, int(slot), out)),
                              self.nbar.get(slot, 0))

    def set_stage_cells(self, slot, cells):
        """Stage cells of one instance (interior edge cells j*nx+i whose outward ghosts hold the stage)."""
        sk = np.ascontiguousarray(np.asarray(cells, dtype=np.int32))
        self._chk(lib().fs_pool_set_stage_cells(self._p, int(slot), sk.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                                len(sk)))

    def set_stage(self, slot, z):
        """Water level of one instance's stage cells' ghosts (m, absolute)."""
        self._chk(lib().fs_pool_set_stage(self._p, int(slot), float(z)))

    def set_rows(self, slot, rows):
        """Rows per wavefront strip of one instance (the autotuned tile shape)."""
        self._chk(lib().fs_pool_set_rows(self._p, int(slot), int(rows)))

    def remove(self, slot):
        self._chk(lib().fs_pool_remove(self._p, int(slot)))
        self.meta.pop(slot, None)

    def set_state(self, slot, h, hu=None, hv=None, bookInitial=True):
        nx, ny, _, lay = self.meta[slot]
        n = nx * ny
        h = np.asarray(h, dtype=np.float64).reshape(n)
        if lay.mask is not None:
            h = np.where(lay.mask != 0, np.maximum(0, h), 0.0)
        h = np.ascontiguousarray(np.maximum(h, 0).astype(np.float32).astype(np.float64))
        hu = np.zeros(n) if hu is None else np.ascontiguousarray(np.asarray(hu, dtype=np.float64).reshape(n))
        hv = np.zeros(n) if hv is None else np.ascontiguousarray(np.asarray(hv, dtype=np.float64).reshape(n))
        self._chk(lib().fs_pool_set_state(self._p, int(slot), _dp(h), _dp(hu), _dp(hv), 1 if bookInitial else 0))

    def set_rain(self, slot, rate):
        self._chk(lib().fs_pool_set_rain(self._p, int(slot), float(rate)))

    def set_inlet_rates(self, slot, q):
        a = np.ascontiguousarray(np.asarray(q, dtype=np.float64))
        self._chk(lib().fs_pool_set_inlet_rates(self._p, int(slot), _dp(a), len(a)))

    def advance(self, slots, targets, K=64):
        s = np.ascontiguousarray(np.asarray(slots, dtype=np.int32))
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(targets, dtype=np.float64), s.shape))
        clks = (fs_clock * len(s))()
        self._chk(lib().fs_pool_advance(self._p, s.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _dp(t), len(s),
                                        int(K), clks))
        return list(clks)

    def step(self, slots, targets, K):
        s = np.ascontiguousarray(np.asarray(slots, dtype=np.int32))
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(targets, dtype=np.float64), s.shape))
        self._chk(lib().fs_pool_step(self._p, s.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _dp(t), len(s), int(K)))

    def clock(self, slot):
        c = fs_clock()
        self._chk(lib().fs_pool_clock(self._p, int(slot), ctypes.byref(c)))
        return c

    def read_state(self, slot, fields=("h", "hu", "hv", "hmax")):
        nx, ny, _, _ = self.meta[slot]
        n = nx * ny
        arrs = {k: np.zeros(n) for k in fields}
        m = fs_mass()
        self._chk(lib().fs_pool_read_state(self._p, int(slot), _dp(arrs.get("h")), _dp(arrs.get("hu")),
                                           _dp(arrs.get("hv")), _dp(arrs.get("hmax")), ctypes.byref(m)))
        arrs["mass"] = {k: getattr(m, k) for k, _ in fs_mass._fields_}
        return arrs

    def set_timing(self, on):
        self._chk(lib().fs_pool_set_timing(self._p, 1 if on else 0))

    def kernel_ms(self):
        a, b, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        self._chk(lib().fs_pool_kernel_ms(self._p, ctypes.byref(a), ctypes.byref(b), ctypes.byref(n)))
        return a.value, b.value, n.value

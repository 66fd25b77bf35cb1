'use strict';
/**
 * GPUSolver over the MI355X HIP library: the surface solver interface of the reference
 * (src/solver-gpu.js:13-693) for Node, in ES2019 syntax, bound to libfloodsim_hip.so through the
 * N-API addon floodsim_napi.node.
 *
 *   const { GPUSolver } = require('floodsim_amd/js/solver-hip.js');
 *   const s = new GPUSolver(null, { nx, ny, dx, z, manning, mask, bc, lts: 4 });
 *   s.setDepth((x, y, z) => ...); s.rainRate = 50 / 3.6e6;
 *   await s.advance(600); const st = await s.readState();
 *
 * What matches the reference:
 * - constructor options nx, ny, dx, z, manning (number or array), mask, infil (number or array),
 *   rainField, bc {west, east, south, north: 'wall'|'open'|'outflow'}, secondOrder, limiter, cfl,
 *   hdry, hfront, dtDry, g, inlets, lts (true = 3 levels, or 1-4), tiles, implicitFriction, stepFlux,
 *   stageCells, crest, ga {psiD, fmax, perv}, gi {rate, cap, drain} (solver-gpu.js:18-20, 35-47, 125-133,
 *   144-151; solver-cpu.js:94-96);
 * - setDepth(fn(x, y, z)) and setVelocity(fn(x, y)) evaluated at cell centres, depth clamped at 0
 *   and 0 in walls, vol.initial the sum of the stored FP32 depths (solver-gpu.js:301-318);
 * - rainRate (m/s), setRainField, setStage(z) (m, the level the stage cells' ghosts hold),
 *   setCrestLevels(zs), readBarriers() -> {plus, minus, peak, level, load, tPeak, tLevel, tLoad, tOver},
 *   setInletRates (m^3/s, + = surface into the network),
 *   readInlets() -> {vol: Float64Array, h: Float32Array, app}, advance(tTarget, K = 64) with the
 *   'GPU solver made no progress (device error?)' error, frame(tEnd, K, readOut) ->
 *   Float32Array (depth, speed per cell) or null, readState() -> {h, hu, hv, hmax, mass},
 *   readHazard(), tileStats(), destroy(); fields t, steps, lastDt, dtAvg, vol, nx, ny, dx, NX, NY, hdry, zref,
 *   zh, idx(i, j), lts, ltsOff, tiled, bc, ready (solver-gpu.js:550-693).
 *
 * The device argument is accepted for signature compatibility and ignored: the library drives
 * HIP device 0 of the process. Options of the reference outside the shallow-water step
 * (subgrid, sink, sediment, debris, scour, banks, underdrains, lines) raise. implicitFriction and stepFlux (F26) are taken
 * as the reference takes them (stepFlux has no effect under local time stepping).
 *
 * Calls on one solver run one at a time: asynchronous calls queue behind each other, and the
 * synchronous setters wait for the queue when it is busy, so they land between the calls they
 * were made between.
 */
const path = require('path');

const addon = require(path.join(__dirname, 'floodsim_napi.node'));
const { def, f64, sidesOf, paramsOf, layersOf, crestColumns, barrierBooks } = require(path.join(__dirname, 'params.js'));

const GH = 2; // the reference's ghost width, kept for idx(i, j) and zh
const UNSUPPORTED = ['subgrid', 'sink', 'sediment', 'debris', 'scour', 'banks', 'underdrains', 'lines'];
// the named parts of a snapshot (fs_snapshot_layout), as `buf` lists them
const SNAP_PARTS = ['record', 'planes', 'barriers', 'soil', 'rainField'];

class GPUSolver {
  constructor(device, o) {
    for (const k of UNSUPPORTED) {
      if (o[k]) throw new Error('GPUSolver (HIP): option ' + k + ' is not supported');
    }
    this.device = device;
    this.nx = o.nx; this.ny = o.ny; this.dx = o.dx;
    this.NX = o.nx + 2 * GH; this.NY = o.ny + 2 * GH;
    this.g = def(o.g, 9.81);
    this.secondOrder = def(o.secondOrder, true);
    this.cfl = def(o.cfl, this.secondOrder ? 0.48 : 0.9);
    this.hdry = def(o.hdry, 1e-5); this.hfront = def(o.hfront, 1e-3);
    this.limiter = o.limiter === 'vanleer' ? 1 : 0;
    this.bc = sidesOf(o.bc);
    this.dtDry = def(o.dtDry, 1.0);
    this.lts = o.lts === true ? 3 : (o.lts | 0);
    this.ltsOff = null;
    this.tiled = o.tiles !== false;
    // hazard products every step, as the reference always tracks them; hazard: false leaves them out
    this.hazard = o.hazard !== false;
    // F26 options, default off (solver-gpu.js:43)
    this.implicitFriction = !!o.implicitFriction; this.stepFlux = !!o.stepFlux;
    const n = o.nx * o.ny;
    const params = paramsOf(o, true);
    const layers = layersOf(o);
    const z = layers[0], rf = layers[5];
    let zmin = Infinity;
    for (let k = 0; k < n; k++) if (z[k] < zmin) zmin = z[k];
    this.zref = zmin;
    this._z = z;
    this._mask = layers[3];
    this._rf = rf;
    this.nInlets = o.inlets ? o.inlets.length : 0;
    const ga = o.ga ? [f64(o.ga.psiD, n, 'ga.psiD'), f64(o.ga.fmax, n, 'ga.fmax'), f64(o.ga.perv, n, 'ga.perv')] : null;
    const gi = o.gi ? [f64(o.gi.rate, n, 'gi.rate'), f64(o.gi.cap, n, 'gi.cap'), f64(o.gi.drain, n, 'gi.drain')] : null;
    this.stage = -1e9;
    // flood barriers: faces {i, j, axis, z, C, ground} (solver-gpu.js:18-20, 144-151)
    this.crest = (o.crest || []).slice();
    this.nBar = this.crest.length;
    // The device objects: an error here (the library, the device, its memory, a layer the grid cannot hold)
    // does not throw from the constructor; `ready` rejects with it, as the reference's does when its pipelines
    // fail (solver-gpu.js:287), and every later call on the solver fails with the same error.
    this._h = null;
    this._buf = null;
    this._createError = null;
    try {
      this._h = addon.create(params, ...layers);
      if (o.stageCells && o.stageCells.length) addon.setStageCells(this._h, Int32Array.from(o.stageCells));
      // Green-Ampt soil and GI stores per cell (solver-cpu.js:94-96)
      if (ga) addon.setGa(this._h, ga[0], ga[1], ga[2]);
      if (gi) addon.setGi(this._h, gi[0], gi[1], gi[2]);
      if (this.nBar) addon.setCrests(this._h, ...crestColumns(this.crest));
      this._layout();
    } catch (e) {
      if (this._h) addon.destroy(this._h);
      this._h = null;
      this._createError = new Error('GPU pipeline error: ' + (e && e.message ? e.message : e));
    }
    this.t = 0; this.steps = 0; this.lastDt = 0; this.dtAvg = 0;
    this.vol = { rain: 0, initial: 0, exchanged: 0 };
    this._rain = 0;
    let area = 0;
    for (let k = 0; k < n; k++) if (!this._mask || this._mask[k]) area += rf ? rf[k] : 1;
    this.rainArea = area * this.dx * this.dx;
    this._depth = new Float64Array(n);
    this._queue = Promise.resolve();
    this._pending = 0;
    this._zh = null;
    if (this._createError) {
      this.ready = Promise.reject(this._createError);
      this.ready.catch(() => {});   // handled by whoever awaits ready; no unhandled-rejection warning
    } else {
      this.ready = Promise.resolve();
    }
  }

  /** Padded index of interior cell (i, j), as the reference's idx (GH = 2 ghost layers). */
  idx(i, j) { return (j + GH) * this.NX + (i + GH); }

  /** Bed relative to its minimum on the padded grid (edge-replicated ghosts), as the reference's zh. */
  get zh() {
    if (!this._zh) {
      const zh = new Float64Array(this.NX * this.NY);
      for (let j = 0; j < this.NY; j++) {
        for (let i = 0; i < this.NX; i++) {
          const ii = Math.min(Math.max(i - GH, 0), this.nx - 1), jj = Math.min(Math.max(j - GH, 0), this.ny - 1);
          zh[j * this.NX + i] = this._z[jj * this.nx + ii] - this.zref;
        }
      }
      this._zh = zh;
    }
    return this._zh;
  }

  // ---- serialisation of calls on the one handle
  _enqueue(fn) {
    if (this._createError) return Promise.reject(this._createError);
    this._pending++;
    const p = this._queue.then(fn);
    const done = () => { this._pending--; };
    this._queue = p.then(done, done);
    return p;
  }

  _sync(fn) {
    if (this._createError) throw this._createError;
    if (this._pending === 0) { fn(); return; }
    this._enqueue(fn).catch((e) => { this._deferredError = e; });
  }

  _raiseDeferred() {
    if (this._deferredError) { const e = this._deferredError; this._deferredError = null; throw e; }
  }

  _clock(c) {
    const ds = c.steps - this.steps;
    if (ds > 0 && c.t > this.t) this.dtAvg = (c.t - this.t) / ds;
    this.lastDt = c.lastDt > 0 ? c.lastDt : this.lastDt;
    this.t = c.t; this.steps = c.steps;
  }

  // ---- state (solver-gpu.js:301-318)
  setDepth(fn) {
    const nx = this.nx, ny = this.ny, dx = this.dx, d = this._depth;
    let v = 0;
    for (let j = 0; j < ny; j++) {
      for (let i = 0; i < nx; i++) {
        const k = j * nx + i;
        const h = !this._mask || this._mask[k] ? Math.max(0, fn((i + 0.5) * dx, (j + 0.5) * dx, this._z[k])) : 0;
        d[k] = Math.fround(h); v += d[k];
      }
    }
    this.vol.initial = v * dx * dx;
    const hu = new Float64Array(nx * ny), hv = new Float64Array(nx * ny), h = Float64Array.from(d);
    this._sync(() => addon.setState(this._h, h, hu, hv, true));
  }

  setVelocity(fn) {
    const nx = this.nx, ny = this.ny, dx = this.dx, d = this._depth;
    const hu = new Float64Array(nx * ny), hv = new Float64Array(nx * ny);
    for (let j = 0; j < ny; j++) {
      for (let i = 0; i < nx; i++) {
        const k = j * nx + i, uv = fn((i + 0.5) * dx, (j + 0.5) * dx);
        hu[k] = d[k] * uv[0]; hv[k] = d[k] * uv[1];
      }
    }
    const h = Float64Array.from(d);
    this._sync(() => addon.setState(this._h, h, hu, hv, true));
  }

  get rainRate() { return this._rain; }
  set rainRate(r) {
    this._rain = +r;
    if (this._h) this._sync(() => addon.setRain(this._h, this._rain));
  }

  setRainField(f) {
    const a = f64(f, this.nx * this.ny, 'rainField');
    this._sync(() => { addon.setRainField(this._h, a); this._layout(); });
  }

  /** Water level (m, absolute) the stage cells' ghosts hold from the next step on (solver-gpu.js:554). */
  setStage(z) {
    this.stage = +z;
    this._sync(() => addon.setStage(this._h, this.stage));
  }

  /** New crest per barrier face, in the order of o.crest (-Infinity = not there; solver-gpu.js:486-488). */
  setCrestLevels(zs) {
    const a = Float64Array.from(zs);
    if (a.length !== this.nBar) throw new RangeError('setCrestLevels: ' + a.length + ' levels for ' + this.nBar + ' faces');
    this._sync(() => addon.setCrestLevels(this._h, a));
  }

  /** Barrier books per face since the start (solver-gpu.js:490-504). */
  readBarriers() {
    return this._enqueue(() => barrierBooks(addon.readBarriers(this._h, this.nBar), this.nBar));
  }

  /** Depth infiltrated so far per interior cell under Green-Ampt (m per pervious area; 0 without it;
   *  solver-gpu.js:507-512). */
  readInfiltration() {
    return this._enqueue(() => addon.readInfiltration(this._h));
  }

  /** Exchange rates per inlet (m^3/s, + = surface into the network), held until the next call. */
  setInletRates(q) {
    const a = Float64Array.from(q);
    this._sync(() => addon.setInletRates(this._h, a));
  }

  /** Volume exchanged per inlet since the last call (m^3) and the inlet depths (m). */
  readInlets() {
    this._raiseDeferred();
    return this._enqueue(() => addon.readInlets(this._h)).then((r) => {
      const h = Float32Array.from(r.h);
      for (let m = 0; m < r.vol.length; m++) this.vol.exchanged += r.vol[m];
      // app: per inlet, the frontal and side discharge arriving at its cell (m^3/s), the frontal neighbour's
      // depth (m) and its axis (0 x, 1 y), as the last exchange left them (kernels.wgsl.js:620-637)
      return { vol: r.vol, h, app: r.app };
    });
  }

  /** Advance to tTarget in batches of K steps; rejects if the device stops making progress. */
  advance(tTarget, K) {
    this._raiseDeferred();
    const t0 = this.t;
    return this._enqueue(() => addon.advance(this._h, tTarget, K === undefined ? 64 : K)).then((c) => {
      this._clock(c);
      this.vol.rain += this._rain * (this.t - t0) * this.rainArea;
    });
  }

  /** Advance up to tEnd with at most K steps, then return packed (depth, speed) or null. */
  frame(tEnd, K, readOut) {
    this._raiseDeferred();
    const t0 = this.t;
    const ro = readOut === undefined ? true : !!readOut;
    return this._enqueue(() => addon.frame(this._h, tEnd, K === undefined ? 64 : K, ro)).then((r) => {
      this._clock(r.clock);
      this.vol.rain += this._rain * (this.t - t0) * this.rainArea;
      return r.out;
    });
  }

  /** Interior fields (Float64Array h, hu, hv, hmax) and the FP64 mass books. */
  readState() {
    this._raiseDeferred();
    return this._enqueue(() => addon.readState(this._h)).then((r) => {
      this._clock(r.clock);
      return { h: r.h, hu: r.hu, hv: r.hv, hmax: r.hmax, mass: r.mass };
    });
  }

  /** Active tiles of the last step (solver-gpu.js:625-629): {tiled, tiles, update, flux} as tile counts
   *  ({tiled: false, tiles} when every tile steps), where this build's tile is its wavefront strip (LTS: the
   *  8x8 level tile); also the open cells the last step advanced. */
  tileStats() {
    return this._enqueue(() => {
      const s = addon.tileStats(this._h);
      if (!s.tiled) return { tiled: false, tiles: s.tiles, active: s.active };
      return { tiled: true, tiles: s.tiles, update: s.update, flux: s.flux, active: s.active };
    });
  }

  /** Hazard products per interior cell (solver-gpu.js:515-529): dv, force, duration, arrival, vmax,
   *  tWet (Float32Array; NaN = never) and the worst class per standard, cls.{arr, defra, defraUrban}. */
  readHazard() {
    this._raiseDeferred();
    return this._enqueue(() => addon.readHazard(this._h));
  }

  /** Everything the solver holds on the device, with its clock and books (checkpoint.js:94-113). */
  snapshot() {
    this._raiseDeferred();
    const meta = { t: this.t, steps: this.steps, lastDt: this.lastDt, dtAvg: this.dtAvg, vol: Object.assign({}, this.vol),
      rainRate: this._rain };
    return this._enqueue(() => addon.snapshot(this._h)).then((bytes) => ({ bytes, meta }));
  }

  /** The device state as named buffers with their sizes in bytes, {record, planes, barriers?, soil?, rainField?}
   *  ({name: {size}}, the parts a snapshot holds; the reference's solver.buf, solver-gpu.js:173-299, which
   *  gpuSnapshot / gpuRestore walk, checkpoint.js:96-112). */
  get buf() {
    if (this._createError) throw this._createError;
    return this._buf;
  }

  _layout() {
    const sizes = addon.snapshotLayout(this._h), b = {};
    SNAP_PARTS.forEach((k, q) => { if (sizes[q] > 0) b[k] = { size: sizes[q] }; });
    this._buf = b;
    return b;
  }

  /** The named parts of a snapshot, {name: Uint8Array} in the order of `buf` (checkpoint.js:96-104). */
  snapshotParts() {
    this._raiseDeferred();
    return this._enqueue(() => {
      const buf = this._layout();
      return addon.snapshot(this._h).then((bytes) => {
        const out = {};
        let off = 0;
        for (const k of Object.keys(buf)) {
          out[k] = bytes.slice(off, off + buf[k].size);
          off += buf[k].size;
        }
        return out;
      });
    });
  }

  /** Put back named parts ({name: Uint8Array}), each checked against the size of the buffer of that name here
   *  (checkpoint.js:106-112); every part this solver holds must be given. The device clock and books come back
   *  with them; the JS fields t, steps, lastDt and dtAvg follow the device clock. */
  restoreParts(parts) {
    const buf = this.buf, names = Object.keys(buf);
    let total = 0;
    for (const k of Object.keys(parts)) {
      const b = buf[k];
      if (!b) continue;
      const n = parts[k].byteLength;
      if (b.size !== n) {
        throw new Error('checkpoint: solver buffer ' + k + ' is ' + b.size + ' bytes here, ' + n +
          ' in the checkpoint (a different grid or options)');
      }
    }
    for (const k of names) {
      if (!parts[k]) throw new Error('checkpoint: solver buffer ' + k + ' is not in the checkpoint');
      total += buf[k].size;
    }
    const all = new Uint8Array(total);
    let off = 0;
    for (const k of names) {
      const p = parts[k];
      all.set(p instanceof Uint8Array ? p : new Uint8Array(p.buffer, p.byteOffset, p.byteLength), off);
      off += buf[k].size;
    }
    this._raiseDeferred();
    return this._enqueue(() => addon.restore(this._h, all)).then(() => {
      const c = addon.clock(this._h);
      this.t = c.t; this.steps = c.steps; this.lastDt = c.lastDt; this.dtAvg = c.dtAvg;
    });
  }

  /** Put back a snapshot of this solver, or of one built from the same model. */
  restore(snap) {
    this._raiseDeferred();
    return this._enqueue(() => addon.restore(this._h, snap.bytes)).then(() => {
      const m = snap.meta;
      this.t = m.t; this.steps = m.steps; this.lastDt = m.lastDt; this.dtAvg = m.dtAvg;
      this.vol = Object.assign({}, m.vol); this._rain = m.rainRate;
    });
  }

  /** Test switch (no counterpart in the reference): with on = true every advance runs no steps, so advance
   *  rejects with the no-progress error after its progress checks; false restores stepping. */
  _debugNoSteps(on) { this._sync(() => addon.debugNoSteps(this._h, !!on)); }

  /** LTS levels of the last cycle per 8x8 tile: {levels: Uint8Array, tx, ty}. */
  levels() { return this._enqueue(() => addon.levels(this._h)); }

  destroy() {
    if (!this._h) return;
    const h = this._h;
    this._h = null;
    if (this._pending === 0) addon.destroy(h);
    else this._queue = this._queue.then(() => addon.destroy(h));
  }
}

module.exports = { GPUSolver, version: addon.version };

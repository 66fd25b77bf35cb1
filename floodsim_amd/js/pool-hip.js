'use strict';
/**
 * Batched and served solves for Node (ES2019, CommonJS): Pool, Server and the batch runner over the
 * fs_pool_* calls of libfloodsim_hip.so (include/floodsim_hip.h), through the N-API addon.
 *
 *   const { Pool, Server, Request, runBatched } = require('floodsim_amd/js/pool-hip.js');
 *   const P = new Pool(0);                               // HBM budget in bytes (0: the device's free memory)
 *   const s = P.add({ nx, ny, dx, z, mask, manning, bc }); // a slot, or null when the budget is spent
 *   P.setState(s, new Float64Array(nx * ny)); P.setRain(s, 40 / 3.6e6);
 *   await P.advance([s], [600]); const st = await P.readState(s);
 *
 * The reference has no counterpart: it runs one solve at a time (src/batch.js:1-4, tools/batch.mjs). The
 * Python forms are floodsim_amd/solver.py (Pool), batch.py (runBatched, tuneRows) and serve.py (Server); both
 * drive the same C calls, so the same inputs give the same bits through either. A pool advances all the
 * slots it is given with one launch sequence, each with its own dt, clock and FP64 books, and an instance's
 * results do not depend on what else is in the pool or on the strip height (setRows).
 *
 * Options of add() are GPUSolver's (solver-hip.js) with the Python Pool's defaults (hazard products off).
 * Calls on one pool run one at a time: advance, step and readState return Promises and queue behind each
 * other; the synchronous calls (setState, setRain, add, remove, clock, ...) must not be made while one of
 * those is in flight (the addon rejects them as busy): await it first.
 */
const path = require('path');

const addon = require(path.join(__dirname, 'floodsim_napi.node'));
const { f64, paramsOf, layersOf, crestColumns, barrierBooks } = require(path.join(__dirname, 'params.js'));

class Pool {
  constructor(budgetBytes) {
    this._p = addon.poolCreate(budgetBytes || 0);
    this.meta = new Map();   // slot -> {nx, ny, mask, nBar}
    this._queue = Promise.resolve();
    this._pending = 0;
  }

  /** Bytes one instance of nx x ny takes in a pool (its device allocation). */
  static bytesNeeded(nx, ny, infil, rainField) {
    return addon.poolBytesNeeded({ nx, ny, dx: 1 }, !!infil, !!rainField);
  }

  bytesUsed() { return addon.poolBytesUsed(this._p); }

  _enqueue(fn) {
    this._pending++;
    const p = this._queue.then(fn);
    const done = () => { this._pending--; };
    this._queue = p.then(done, done);
    return p;
  }

  _idle(what) {
    if (this._pending) throw new Error('Pool.' + what + ': a call on this pool is still in flight (await it first)');
  }

  /** Admit one instance (GPUSolver's options plus stageCells, crest, ga, gi); null when the budget is spent. */
  add(o) {
    this._idle('add');
    const slot = addon.poolAdd(this._p, paramsOf(o, false), ...layersOf(o));
    if (slot < 0) return null;
    const n = o.nx * o.ny;
    const mask = o.mask ? (o.mask instanceof Uint8Array ? o.mask : Uint8Array.from(o.mask)) : null;
    const crest = o.crest || [];
    this.meta.set(slot, { nx: o.nx, ny: o.ny, mask, nBar: crest.length });
    try {
      if (o.stageCells && o.stageCells.length) addon.poolSetStageCells(this._p, slot, Int32Array.from(o.stageCells));
      if (crest.length) addon.poolSetCrests(this._p, slot, ...crestColumns(crest));
      if (o.ga) addon.poolSetGa(this._p, slot, f64(o.ga.psiD, n, 'ga.psiD'), f64(o.ga.fmax, n, 'ga.fmax'), f64(o.ga.perv, n, 'ga.perv'));
      if (o.gi) addon.poolSetGi(this._p, slot, f64(o.gi.rate, n, 'gi.rate'), f64(o.gi.cap, n, 'gi.cap'), f64(o.gi.drain, n, 'gi.drain'));
    } catch (e) {
      this.remove(slot);
      throw e;
    }
    return slot;
  }

  remove(slot) {
    this._idle('remove');
    addon.poolRemove(this._p, slot);
    this.meta.delete(slot);
  }

  /** Rows per wavefront strip of one instance (the autotuned tile shape; results do not depend on it). */
  setRows(slot, rows) { this._idle('setRows'); addon.poolSetRows(this._p, slot, rows); }

  /** Depth (and discharges) per cell: depth clamped at 0 and 0 in walls, stored as FP32, as GPUSolver.setDepth. */
  setState(slot, h, hu, hv, bookInitial) {
    this._idle('setState');
    const m = this.meta.get(slot);
    if (!m) throw new RangeError('Pool.setState: no instance in slot ' + slot);
    const n = m.nx * m.ny;
    const d = new Float64Array(n);
    for (let k = 0; k < n; k++) d[k] = !m.mask || m.mask[k] ? Math.fround(Math.max(0, h[k])) : 0;
    addon.poolSetState(this._p, slot, d, hu ? f64(hu, n, 'hu') : new Float64Array(n), hv ? f64(hv, n, 'hv') : new Float64Array(n),
      bookInitial === undefined ? true : !!bookInitial);
  }

  setRain(slot, rate) { this._idle('setRain'); addon.poolSetRain(this._p, slot, +rate); }
  setInletRates(slot, q) { this._idle('setInletRates'); addon.poolSetInletRates(this._p, slot, Float64Array.from(q)); }
  setStageCells(slot, cells) { this._idle('setStageCells'); addon.poolSetStageCells(this._p, slot, Int32Array.from(cells)); }
  setStage(slot, z) { this._idle('setStage'); addon.poolSetStage(this._p, slot, +z); }

  setCrestLevels(slot, zs) {
    this._idle('setCrestLevels');
    const m = this.meta.get(slot), a = Float64Array.from(zs);
    if (!m || a.length !== m.nBar) throw new RangeError('Pool.setCrestLevels: ' + a.length + ' levels for ' + (m ? m.nBar : 0) + ' faces');
    addon.poolSetCrestLevels(this._p, slot, a);
  }

  readBarriers(slot) {
    this._idle('readBarriers');
    const m = this.meta.get(slot), n = m ? m.nBar : 0;
    return barrierBooks(addon.poolReadBarriers(this._p, slot, n), n);
  }

  /** Advance the listed slots together until each reaches its own target: Promise<clock per slot>. */
  advance(slots, targets, K) {
    const s = Int32Array.from(slots), t = broadcast(targets, s.length);
    return this._enqueue(() => addon.poolAdvance(this._p, s, t, K === undefined ? 64 : K));
  }

  /** Exactly K steps of every listed slot (stopping at its target): Promise<undefined>. */
  step(slots, targets, K) {
    const s = Int32Array.from(slots), t = broadcast(targets, s.length);
    return this._enqueue(() => addon.poolStep(this._p, s, t, K));
  }

  /** {t, lastDt, dtAvg, steps, cellUpdates} of one slot (waits for the device). */
  clock(slot) { this._idle('clock'); return addon.poolClock(this._p, slot); }

  /** Interior fields of one slot (Float64Array h, hu, hv, hmax), its FP64 books and clock. */
  readState(slot) { return this._enqueue(() => addon.poolReadState(this._p, slot)); }

  setTiming(on) { this._idle('setTiming'); addon.poolSetTiming(this._p, !!on); }
  kernelMs() { this._idle('kernelMs'); return addon.poolKernelMs(this._p); }

  destroy() {
    if (!this._p) return;
    const p = this._p;
    this._p = null;
    if (this._pending === 0) addon.poolDestroy(p);
    else this._queue = this._queue.then(() => addon.poolDestroy(p));
  }
}

function broadcast(t, n) {
  if (typeof t === 'number') return new Float64Array(n).fill(t);
  const a = Float64Array.from(t);
  if (a.length !== n) throw new RangeError('targets: ' + a.length + ' for ' + n + ' slots');
  return a;
}

function seconds() {
  const t = process.hrtime();
  return t[0] + 1e-9 * t[1];
}

// ---- batched solves (floodsim_amd/batch.py)

const ROW_CANDIDATES = [16, 32, 64, 128];

/** Admit every instance ({nx, ny, dx, z, mask, manning, bc, rain}) dry, with its rain rate: the slots. */
function addAll(pool, insts, rows) {
  return insts.map((inst) => {
    const s = pool.add(Object.assign({}, inst, { rows: rows || 0 }));
    if (s === null) throw new Error('pool budget too small for the batch');
    pool.setState(s, new Float64Array(inst.nx * inst.ny));
    pool.setRain(s, inst.rain || 0);
    return s;
  });
}

/** Autotuner: time `steps` real steps of the solve per candidate strip height and keep the fastest (the choice
 *  changes no result). Promise<{rows, timing: {rows: ms per step}}>. */
async function tuneRows(pool, slots, targets, candidates, steps, repeats) {
  const cand = candidates || ROW_CANDIDATES, ns = steps || 4, nr = repeats || 2;
  const timing = {};
  for (const r of cand) {
    for (const s of slots) pool.setRows(s, r);
    await pool.step(slots, targets, 1);             // rebuild the group and capture its graph
    let best = Infinity;
    for (let q = 0; q < nr; q++) {
      pool.clock(slots[0]);                          // synchronise
      const t0 = seconds();
      await pool.step(slots, targets, ns);
      pool.clock(slots[0]);
      best = Math.min(best, (seconds() - t0) / ns);
    }
    timing[r] = best * 1e3;
  }
  let rows = cand[0];
  for (const r of cand) if (timing[r] < timing[rows]) rows = r;
  for (const s of slots) pool.setRows(s, rows);
  return { rows, timing };
}

/** All instances in one pool, advanced together to tEnd: Promise<{wallS, rows, states, steps, cellUpdates}>. */
async function runBatched(insts, tEnd, o) {
  const opt = o || {};
  const P = new Pool(opt.budget || 0);
  try {
    const slots = addAll(P, insts);
    const targets = slots.map(() => tEnd);
    const t0 = seconds();
    const tuned = opt.tune === false ? { rows: opt.rows || 64, timing: {} } : await tuneRows(P, slots, targets);
    if (opt.tune === false) for (const s of slots) P.setRows(s, tuned.rows);
    await P.advance(slots, targets, opt.K || 64);
    P.clock(slots[0]);
    const wallS = seconds() - t0;
    const clocks = slots.map((s) => P.clock(s));
    const states = [];
    if (opt.read !== false) for (const s of slots) states.push(await P.readState(s));
    return { wallS, rows: tuned.rows, tuneMs: tuned.timing, states, steps: clocks.map((c) => c.steps),
      cellUpdates: clocks.reduce((a, c) => a + c.cellUpdates, 0) };
  } finally {
    P.destroy();
  }
}

// ---- served solves (floodsim_amd/serve.py)

class Request {
  /** A solve that arrives at `arrival` (s, on the server's clock): layers {nx, ny, dx, z, mask, manning, bc},
   *  `duration` simulated seconds, rainFn(t) -> m/s. */
  constructor(id, arrival, layers, duration, rainFn, meta) {
    this.id = id;
    this.arrival = arrival;
    this.layers = layers;
    this.duration = duration;
    this.rainFn = rainFn;
    this.meta = meta || {};
    this.slot = null;
    this.t = 0;
    this.admitted = null;
    this.done = null;
    this.result = null;
  }

  get cells() { return this.layers.nx * this.layers.ny; }
}

/**
 * Continuous batching over one Pool: each step() admits the waiting requests that have arrived while the
 * HBM budget has room, advances every active request by one chunk of its own simulated time (one launch
 * sequence for all), with the rain held at the chunk's mid-time value, and retires those that finish.
 * o: chunkS (60), K (64), rows (64), clock (() -> s; wall clock by default), keep (['hmax']).
 */
class Server {
  constructor(budgetBytes, o) {
    const opt = o || {};
    this.pool = new Pool(budgetBytes);
    this.budget = budgetBytes;
    this.chunk = opt.chunkS === undefined ? 60 : +opt.chunkS;
    this.K = opt.K || 64;
    this.rows = opt.rows || 64;
    this.clock = opt.clock || seconds;
    this.t0 = this.clock();
    this.waiting = [];
    this.active = [];
    this.finished = [];
    this.keep = opt.keep || ['hmax'];
    this.iterations = 0;
    this.cellUpdates = 0;
  }

  now() { return this.clock() - this.t0; }

  submit(req) { this.waiting.push(req); }

  _admit() {
    while (this.waiting.length && this.waiting[0].arrival <= this.now()) {
      const r = this.waiting[0];
      const slot = this.pool.add(Object.assign({}, r.layers, { rows: this.rows }));
      if (slot === null) {
        if (!this.active.length) throw new Error('request ' + r.id + ' (' + r.cells + ' cells) does not fit the HBM budget alone');
        return;
      }
      this.waiting.shift();
      this.pool.setState(slot, new Float64Array(r.cells));
      r.slot = slot;
      r.admitted = this.now();
      this.active.push(r);
    }
  }

  /** One iteration; Promise<false> when nothing was active. */
  async step() {
    this._admit();
    if (!this.active.length) return false;
    const slots = [], targets = [];
    for (const r of this.active) {
      const t1 = Math.min(r.duration, r.t + this.chunk);
      this.pool.setRain(r.slot, r.rainFn(0.5 * (r.t + t1)));
      slots.push(r.slot);
      targets.push(t1);
    }
    const clocks = await this.pool.advance(slots, targets, this.K);
    this.iterations++;
    const still = [];
    for (let q = 0; q < this.active.length; q++) {
      const r = this.active[q], c = clocks[q];
      this.cellUpdates += c.cellUpdates - (r.meta._updates || 0);
      r.meta._updates = c.cellUpdates;
      r.t = targets[q];
      if (r.t >= r.duration) {
        const st = await this.pool.readState(r.slot);
        const res = { steps: c.steps, mass: st.mass };
        for (const k of this.keep) res[k] = st[k];
        r.result = res;
        r.done = this.now();
        this.pool.remove(r.slot);
        r.slot = null;
        this.finished.push(r);
      } else {
        still.push(r);
      }
    }
    this.active = still;
    return true;
  }

  /** Serve a list of requests (arrival times on the server clock) until all are done. */
  async run(requests, idleSleepS) {
    const idle = idleSleepS === undefined ? 0.002 : idleSleepS;
    requests.slice().sort((a, b) => a.arrival - b.arrival).forEach((r) => this.submit(r));
    while (this.waiting.length || this.active.length) {
      if (!(await this.step()) && this.waiting.length) {
        const wait = Math.max(0, Math.min(idle * 50, this.waiting[0].arrival - this.now()));
        await new Promise((res) => setTimeout(res, 1e3 * wait));
      }
    }
    return this.finished;
  }

  stats() {
    const lat = this.finished.map((r) => r.done - r.arrival).sort((a, b) => a - b);
    const span = Math.max.apply(null, this.finished.map((r) => r.done)) - Math.min.apply(null, this.finished.map((r) => r.arrival));
    return { requests: this.finished.length, p50S: percentile(lat, 50), p99S: percentile(lat, 99),
      meanS: lat.reduce((a, b) => a + b, 0) / lat.length, spanS: span,
      solvesPerHour: 3600 * this.finished.length / span, cellUpdatesPerS: this.cellUpdates / span,
      iterations: this.iterations };
  }

  close() { this.pool.destroy(); }
}

/** Linear-interpolation percentile of sorted values (numpy's default). */
function percentile(sorted, p) {
  if (!sorted.length) return NaN;
  const x = (sorted.length - 1) * p / 100, lo = Math.floor(x), hi = Math.min(lo + 1, sorted.length - 1);
  return sorted[lo] + (sorted[hi] - sorted[lo]) * (x - lo);
}

/** The same request through the same chunk schedule in a pool of its own. */
async function solveAlone(req, o) {
  const opt = Object.assign({ chunkS: 60, K: 64, rows: 64, keep: ['hmax'] }, o || {}, { clock: () => 0 });
  const S = new Server(0, opt);
  try {
    const r = new Request(req.id, 0, req.layers, req.duration, req.rainFn, Object.assign({}, req.meta, { _updates: 0 }));
    await S.run([r]);
    return r.result;
  } finally {
    S.close();
  }
}

module.exports = { Pool, Server, Request, addAll, tuneRows, runBatched, solveAlone, ROW_CANDIDATES };

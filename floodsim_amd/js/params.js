'use strict';
/**
 * Options -> the C ABI's fs_params and layer arrays, shared by GPUSolver (solver-hip.js) and Pool
 * (pool-hip.js). Defaults are the reference GPUSolver's (src/solver-gpu.js:39-47): second order,
 * cfl 0.48 (0.9 first order), hdry 1e-5, hfront 1e-3, dtDry 1 s, g 9.81, minmod, walls on every side.
 */
const BC = { wall: 0, open: 1, outflow: 2 };
const SIDES = ['west', 'east', 'south', 'north'];

function def(v, d) { return v === undefined || v === null ? d : v; }

/** A Float64Array of n values from a number (filled), an array or a typed array; null for null/undefined. */
function f64(a, n, what) {
  if (a === undefined || a === null) return null;
  if (typeof a === 'number') return new Float64Array(n).fill(a);
  if (a.length !== n) throw new RangeError(what + ': length ' + a.length + ', expected ' + n);
  return a instanceof Float64Array ? a : Float64Array.from(a);
}

/** The boundary sides with their defaults; throws on an unknown kind. */
function sidesOf(bc) {
  const b = Object.assign({ west: 'wall', east: 'wall', south: 'wall', north: 'wall' }, bc || {});
  for (const side of SIDES) {
    if (BC[b[side]] === undefined) throw new Error('GPUSolver (HIP): bc.' + side + ' = ' + b[side]);
  }
  return b;
}

/** fs_params as the addon reads it. hazardDefault: whether hazard products are kept when o.hazard is unset. */
function paramsOf(o, hazardDefault) {
  const secondOrder = def(o.secondOrder, true);
  const bc = sidesOf(o.bc);
  const hazard = o.hazard === undefined || o.hazard === null ? !!hazardDefault : !!o.hazard;
  return {
    nx: o.nx, ny: o.ny, dx: o.dx, g: def(o.g, 9.81), cfl: def(o.cfl, secondOrder ? 0.48 : 0.9),
    hdry: def(o.hdry, 1e-5), hfront: def(o.hfront, 1e-3), dtDry: def(o.dtDry, 1.0),
    secondOrder: secondOrder ? 1 : 0, limiter: o.limiter === 'vanleer' ? 1 : 0,
    bcWest: BC[bc.west], bcEast: BC[bc.east], bcSouth: BC[bc.south], bcNorth: BC[bc.north],
    lts: o.lts === true ? 3 : (o.lts | 0), rows: o.rows | 0, untiled: o.tiles === false ? 1 : 0,
    hazard: hazard ? 1 : 0, implicitFriction: o.implicitFriction ? 1 : 0, stepFlux: o.stepFlux ? 1 : 0,
  };
}

/** The layer arguments of create / poolAdd: [z, manning|null, manningScalar, mask|null, infil|null,
 *  rainField|null, inlets|null]. */
function layersOf(o) {
  const n = o.nx * o.ny;
  const z = f64(o.z, n, 'z');
  if (!z) throw new TypeError('z is required');
  const manArr = typeof o.manning === 'number' || !o.manning ? null : f64(o.manning, n, 'manning');
  const manS = typeof o.manning === 'number' ? o.manning : 0;
  const mask = o.mask ? (o.mask instanceof Uint8Array ? o.mask : Uint8Array.from(o.mask)) : null;
  if (mask && mask.length !== n) throw new RangeError('mask: length ' + mask.length + ', expected ' + n);
  const infil = f64(o.infil, n, 'infil');
  const rf = f64(o.rainField, n, 'rainField');
  const inlets = o.inlets && o.inlets.length ? Int32Array.from(o.inlets) : null;
  return [z, manArr, manS, mask, infil, rf, inlets];
}

/** setCrests columns from the reference's barrier faces [{i, j, axis, z, C, ground}] (solver-gpu.js:18-20). */
function crestColumns(crest) {
  const col = (k, d, T) => T.from(crest.map((f) => (f[k] === undefined || f[k] === null ? d : f[k])));
  return [col('i', 0, Int32Array), col('j', 0, Int32Array), col('axis', 0, Int32Array),
    col('z', -Infinity, Float64Array), col('C', 1.6, Float64Array), col('ground', -Infinity, Float64Array)];
}

const BARRIER_BOOKS = ['plus', 'minus', 'peak', 'level', 'load', 'tPeak', 'tLevel', 'tLoad', 'tOver'];

/** readBarriers' flat Float64Array(9 n) as {plus, minus, ...} per face. */
function barrierBooks(a, n) {
  const o = {};
  BARRIER_BOOKS.forEach((k, q) => {
    o[k] = new Float64Array(n);
    for (let m = 0; m < n; m++) o[k][m] = a[9 * m + q];
  });
  return o;
}

module.exports = { BC, def, f64, sidesOf, paramsOf, layersOf, crestColumns, barrierBooks };

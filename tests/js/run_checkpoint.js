'use strict';
// A run checkpointed and resumed through the JavaScript checkpoint adapter (floodsim_amd/js/checkpoint-hip.js),
// written as the reference's app uses src/checkpoint.js (app/map.js:2616-2660): gpuSnapshot + solverState of a
// running solver, gpuRestore + applySolverState into a new solver built from the same options, then both run on.
// Prints one JSON line: the resumed run's fields beside the uninterrupted run's, the part names and sizes, and
// the errors of a restore into another grid and of a snapshot with a part left out.
const fs = require('fs');
const path = require('path');
const js = path.join(__dirname, '..', '..', 'floodsim_amd', 'js');
const { GPUSolver } = require(path.join(js, 'solver-hip.js'));
const CK = require(path.join(js, 'checkpoint-hip.js'));

function options(c) {
  return { nx: c.nx, ny: c.ny, dx: c.dx, z: Float64Array.from(c.z), mask: Uint8Array.from(c.mask),
    manning: Float64Array.from(c.manning), bc: c.bc, lts: c.lts, infil: Float64Array.from(c.infil),
    stageCells: c.stageCells, crest: c.crest, ga: c.ga, gi: c.gi, hazard: false };
}

async function build(c) {
  const s = new GPUSolver(null, options(c));
  await s.ready;
  s.setStage(c.stage);
  s.setRainField(Float64Array.from(c.rainField));   // given after creation: a part of its own
  return s;
}

function fields(st) {
  return { h: Array.from(st.h), hu: Array.from(st.hu), hv: Array.from(st.hv), hmax: Array.from(st.hmax), mass: st.mass };
}

async function main() {
  const c = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
  const A = await build(c);
  A.setDepth(() => 0);
  A.rainRate = c.rain;
  for (const t of c.frames) await A.advance(t);
  const snap = await CK.gpuSnapshot(A);
  const state = JSON.parse(JSON.stringify(CK.solverState(A)));   // as a checkpoint file would carry it
  const names = Object.keys(snap);
  const sizes = names.map((k) => snap[k].byteLength), bufSizes = names.map((k) => A.buf[k].size);
  const tSnap = A.t, stepsSnap = A.steps;
  await A.advance(c.tEnd);
  const a = await A.readState();
  const stepsA = A.steps;
  const bars = await A.readBarriers();
  const B = await build(c);
  B.setRainField(new Float64Array(c.nx * c.ny).fill(1));       // another field: the snapshot's must come back
  await CK.gpuRestore(B, snap);
  CK.applySolverState(B, state);
  const tB = B.t, stepsB = B.steps, rainB = B.rainRate;
  await B.advance(c.tEnd);
  const b = await B.readState();
  const stepsBEnd = B.steps;
  const barsB = await B.readBarriers();
  // a solver of another grid: the size check throws before anything is written
  const o2 = { nx: c.nx + 2, ny: c.ny, dx: c.dx, z: new Float64Array((c.nx + 2) * c.ny), manning: 0.03, hazard: false };
  const C = new GPUSolver(null, o2);
  await C.ready;
  let errGrid = '', errMissing = '';
  try { CK.gpuRestore(C, snap); } catch (e) { errGrid = e.message; }
  const partial = Object.assign({}, snap);
  delete partial.soil;
  try { CK.gpuRestore(B, partial); } catch (e) { errMissing = e.message; }
  const sameBars = Object.keys(bars).every((k) => bars[k].every((v, m) => Object.is(v, barsB[k][m])));   // (NaN: never)
  A.destroy(); B.destroy(); C.destroy();
  process.stdout.write(JSON.stringify({ names, sizes, bufSizes, tSnap, stepsSnap, tB, stepsB, rainB,
    stepsA, stepsBEnd, a: fields(a), b: fields(b), sameBars,
    volA: A.vol, volB: B.vol, errGrid, errMissing }) + '\n');
}

main().catch((e) => { process.stderr.write(String(e && e.stack || e) + '\n'); process.exit(1); });

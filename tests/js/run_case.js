'use strict';
// Drives one case through the JavaScript GPUSolver (floodsim_amd/js/solver-hip.js) and prints the
// result as one JSON line. The case file holds the constructor options plus the initial depth
// (per cell, row-major), rain, inlet rates, the frame times to advance through and, optionally, the
// stage to hold at the stage cells during each frame.
const fs = require('fs');
const path = require('path');
const { GPUSolver } = require(path.join(__dirname, '..', '..', 'floodsim_amd', 'js', 'solver-hip.js'));

async function main() {
  const c = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
  const o = {
    nx: c.nx, ny: c.ny, dx: c.dx, z: Float64Array.from(c.z), manning: c.manning,
    mask: c.mask ? Uint8Array.from(c.mask) : undefined, bc: c.bc, lts: c.lts || 0,
    inlets: c.inlets && c.inlets.length ? c.inlets : undefined,
    stageCells: c.stageCells, implicitFriction: !!c.implicitFriction, stepFlux: !!c.stepFlux,
    hazard: c.hazard !== false,
  };
  const s = new GPUSolver(null, o);
  await s.ready;
  const depth = c.depth;
  s.setDepth((x, y) => depth[Math.floor(y / c.dx) * c.nx + Math.floor(x / c.dx)]);
  s.rainRate = c.rain || 0;
  if (c.rates) s.setInletRates(c.rates);
  let frameSum = 0, inletVol = 0;
  for (let f = 0; f < c.frames.length; f++) {
    if (c.stageLevels) s.setStage(c.stageLevels[f]);
    const out = await s.frame(c.frames[f], 1 << 20, true);
    for (let k = 0; k < out.length; k += 2) frameSum += out[k];
  }
  if (s.nInlets) { const r = await s.readInlets(); for (const v of r.vol) inletVol += v; }
  const st = await s.readState();
  const tiles = await s.tileStats();
  let hzCells = 0, hzForce = 0;
  if (s.hazard) {
    const hz = await s.readHazard();
    for (let k = 0; k < hz.cls.arr.length; k++) if (hz.cls.arr[k] > 0) hzCells++;
    hzForce = Math.max.apply(null, Array.from(hz.force));
  }
  const t1 = s.t, steps1 = s.steps, lastDt1 = s.lastDt, volRain1 = s.vol.rain;
  // snapshot, run on, restore, run the same span again: the same bits
  const snap = await s.snapshot();
  const tEnd = s.t + 3;
  await s.advance(tEnd);
  const a = await s.readState();
  await s.restore(snap);
  const tBack = s.t;
  await s.advance(tEnd);
  const b = await s.readState();
  let same = a.h.length === b.h.length;
  for (let k = 0; k < a.h.length && same; k++) same = a.h[k] === b.h[k] && a.hu[k] === b.hu[k] && a.hv[k] === b.hv[k];
  s.destroy();
  process.stdout.write(JSON.stringify({
    t: t1, steps: steps1, lastDt: lastDt1, h: Array.from(st.h), hu: Array.from(st.hu), hv: Array.from(st.hv),
    hmax: Array.from(st.hmax), mass: st.mass, frameSum, inletVol, volRain: volRain1, tiles,
    snapshotBytes: snap.bytes.length, hazardCells: hzCells, hazardForceMax: hzForce, restoredT: tBack, restoreSame: same,
  }) + '\n');
}

main().catch((e) => { process.stderr.write(String(e && e.stack || e) + '\n'); process.exit(1); });

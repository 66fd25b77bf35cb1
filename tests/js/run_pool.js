'use strict';
// Drives batched and served solves through the JavaScript Pool and Server (floodsim_amd/js/pool-hip.js) and
// prints the results as one JSON line. The case file holds
//   batch: {instances: [{nx, ny, dx, z, mask, manning, bc, rain, rows, rows2}], t1, T}: every instance dry in one
//          pool, advanced together to t1, its strip height changed to rows2, then on to T;
//   serve: {budget, chunk, requests: [{id, arrival, duration, layers, rain: [[t, rate], ...]}]}: a Server on a
//          virtual clock that moves 1 s per iteration, the rain of each chunk looked up at its mid time.
const fs = require('fs');
const path = require('path');
const { Pool, Server, Request } = require(path.join(__dirname, '..', '..', 'floodsim_amd', 'js', 'pool-hip.js'));

function layers(o) {
  return { nx: o.nx, ny: o.ny, dx: o.dx, z: Float64Array.from(o.z), mask: Uint8Array.from(o.mask),
    manning: typeof o.manning === 'number' ? o.manning : Float64Array.from(o.manning), bc: o.bc };
}

function rainTable(pairs) {
  const m = new Map(pairs.map((p) => [p[0], p[1]]));
  return (t) => {
    const v = m.get(t);
    if (v === undefined) throw new Error('no rain rate for mid time ' + t);
    return v;
  };
}

async function main() {
  const c = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
  const out = {};
  if (c.batch) {
    const P = new Pool(0);
    const insts = c.batch.instances;
    const slots = insts.map((inst) => {
      const s = P.add(Object.assign(layers(inst), { rows: inst.rows || 0 }));
      P.setState(s, new Float64Array(inst.nx * inst.ny));
      P.setRain(s, inst.rain);
      return s;
    });
    await P.advance(slots, c.batch.t1);
    slots.forEach((s, q) => { if (insts[q].rows2) P.setRows(s, insts[q].rows2); });
    const clocks = await P.advance(slots, slots.map(() => c.batch.T));
    const states = [];
    for (const s of slots) {
      const st = await P.readState(s);
      states.push({ h: Array.from(st.h), hu: Array.from(st.hu), hv: Array.from(st.hv), hmax: Array.from(st.hmax),
        mass: st.mass });
    }
    out.batch = { steps: clocks.map((k) => k.steps), t: clocks.map((k) => k.t), bytesUsed: P.bytesUsed(), states };
    P.destroy();
  }
  if (c.serve) {
    let tick = 0;
    const S = new Server(c.serve.budget, { chunkS: c.serve.chunk, clock: () => tick });
    const reqs = c.serve.requests.map((r) => new Request(r.id, r.arrival, layers(r.layers), r.duration, rainTable(r.rain)));
    for (const r of reqs) S.submit(r);
    let peak = 0, most = 0;
    while (S.waiting.length || S.active.length) {
      await S.step();
      peak = Math.max(peak, S.pool.bytesUsed());
      most = Math.max(most, S.active.length);
      tick += 1;
    }
    const st = S.stats();
    S.close();
    out.serve = { peak, most, iterations: st.iterations, requests: reqs.map((r) => ({ id: r.id, admitted: r.admitted,
      done: r.done, steps: r.result.steps, hmax: Array.from(r.result.hmax), mass: r.result.mass })) };
  }
  process.stdout.write(JSON.stringify(out) + '\n');
}

main().catch((e) => { process.stderr.write(String(e && e.stack || e) + '\n'); process.exit(1); });

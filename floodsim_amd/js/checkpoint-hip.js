'use strict';
/**
 * Run checkpoints of the HIP GPUSolver with the reference's checkpoint interface (src/checkpoint.js:94-114):
 *
 *   const snap = await gpuSnapshot(solver);          // {record, planes, barriers?, soil?, rainField?: Uint8Array}
 *   const js = solverState(solver);                  // the solver's own JS state (clock, books, rain, stage)
 *   ...                                              // later, a solver built from the same model and options:
 *   gpuRestore(solver2, snap); applySolverState(solver2, js);
 *   await solver2.advance(t);                        // continues bit for bit where the first one was
 *
 * gpuSnapshot returns every device buffer of the solver by name, as the reference's does for its storage
 * buffers (checkpoint.js:96-104); the names and sizes are `solver.buf`'s. gpuRestore checks each named part
 * against the size of the buffer of that name and throws the reference's error on a mismatch, before writing
 * anything (checkpoint.js:106-112); a part this solver holds that the snapshot lacks is an error too (the
 * reference leaves such a buffer as it is; here the parts go back as one copy). The write is queued on the
 * solver like its other calls, so a call made after gpuRestore runs after the state is back; the returned
 * Promise resolves then. solverState / applySolverState carry the JS fields a resumed solver needs (the
 * reference captures the solver object's fields, checkpoint.js:114).
 */
const JS_FIELDS = ['t', 'steps', 'lastDt', 'dtAvg', 'stage'];

/** Every device buffer of the solver, read after its queued calls: Promise<{[name]: Uint8Array}>. */
function gpuSnapshot(s) { return s.snapshotParts(); }

/** Write a snapshot back into a solver built with the same grid and options (buffer sizes must match). */
function gpuRestore(s, snap) { return s.restoreParts(snap); }

/** The solver's JS state (clock, volume books, rain rate, stage), JSON-able. */
function solverState(s) {
  const o = { vol: Object.assign({}, s.vol), rainRate: s.rainRate };
  for (const k of JS_FIELDS) o[k] = s[k];
  return o;
}

/** Put a solverState back into a solver (the device state comes back through gpuRestore). */
function applySolverState(s, st) {
  for (const k of JS_FIELDS) if (st[k] !== undefined) s[k] = st[k];
  if (st.vol) s.vol = Object.assign({}, st.vol);
  if (st.rainRate !== undefined) s._rain = +st.rainRate;   // the device's rate is in the snapshot's record
  return s;
}

module.exports = { gpuSnapshot, gpuRestore, solverState, applySolverState };

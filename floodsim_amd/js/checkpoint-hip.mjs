// ES-module face of checkpoint-hip.js, for importers written like the reference's src/checkpoint.js callers.
import cjs from './checkpoint-hip.js';

export const gpuSnapshot = cjs.gpuSnapshot;
export const gpuRestore = cjs.gpuRestore;
export const solverState = cjs.solverState;
export const applySolverState = cjs.applySolverState;

// ES-module face of pool-hip.js.
import cjs from './pool-hip.js';

export const Pool = cjs.Pool;
export const Server = cjs.Server;
export const Request = cjs.Request;
export const addAll = cjs.addAll;
export const tuneRows = cjs.tuneRows;
export const runBatched = cjs.runBatched;
export const solveAlone = cjs.solveAlone;

// ES-module face of solver-hip.js, for importers written like the reference's src/*.js modules.
import cjs from './solver-hip.js';

export const GPUSolver = cjs.GPUSolver;
export const version = cjs.version;

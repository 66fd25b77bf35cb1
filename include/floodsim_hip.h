/*
 * floodsim_hip: MI355X (gfx950) shallow-water step behind a C ABI.
 *
 * The surface solver of the reference (src/solver-gpu.js, class GPUSolver) is a
 * host object that owns its device buffers, is built from the model layers,
 * advanced in time and read back. This header is the C-ABI replacement of that
 * object; the N-API addon (floodsim_amd/csrc/napi_addon.c) and the JavaScript
 * class (floodsim_amd/js/solver-hip.js) bind it one to one.
 *
 * Conventions: every array is interior-only, nx*ny, row-major (k = j*nx + i), in
 * SI units (m, m/s, m^3); the bed is absolute (the library stores it relative to
 * its minimum, as solver-gpu.js:52-53, 65 does). Every call returns 0 on success
 * or a negative FS_E_* code; fs_last_error() gives the message. Calls on one
 * handle must not overlap (the JS side serialises them).
 */
#ifndef FLOODSIM_HIP_H
#define FLOODSIM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FS_OK 0
#define FS_E_ARG (-1)      /* bad argument */
#define FS_E_HIP (-2)      /* a HIP runtime call failed */
#define FS_E_BUDGET (-3)   /* admission refused: the pool's HBM budget is spent */
#define FS_E_STALL (-4)    /* advance made no progress (solver-gpu.js:601) */
#define FS_E_STATE (-5)    /* call not valid in this state */

enum { FS_BC_WALL = 0, FS_BC_OPEN = 1, FS_BC_OUTFLOW = 2 };

/* Constructor options of GPUSolver (solver-gpu.js:35-158; defaults at :39-47). */
typedef struct fs_params {
  int32_t nx, ny;      /* interior cells */
  double dx;           /* cell size, m */
  double g;            /* 9.81 */
  double cfl;          /* 0.48 second order, 0.9 first order */
  double hdry;         /* 1e-5 (GPU default, solver-gpu.js:42) */
  double hfront;       /* 1e-3 */
  double dtDry;        /* step while all is dry, s (1.0) */
  int32_t secondOrder; /* 1 = MUSCL + SSP-RK2 */
  int32_t limiter;     /* 0 minmod, 1 van Leer */
  int32_t bc[4];       /* west, east, south, north: FS_BC_* */
  int32_t lts;         /* local time stepping levels 0..4 (0 = global steps) */
  int32_t rows;        /* rows per wavefront strip (0 = autotuned; local time stepping: a multiple of 8, at most 64) */
  int32_t untiled;     /* 1: advance every strip every step (the reference's tiles: false); 0: skip
                          quiet dry strips (the reference's active tiles, kernels.wgsl.js:639-673) */
  int32_t hazard;      /* 1: track the hazard products every step (hazard/hzRecord,
                          kernels.wgsl.js:415-447, which the reference always runs) */
  int32_t implicitFriction; /* F26 (solver-cpu.js:45, 62; solver-gpu.js:43): Manning friction implicit in the
                          new speed, and the second stage's fluxes see the predictor after friction */
  int32_t stepFlux;    /* F26 (solver-cpu.js:47, 63): a film stepping down a bed step taller than itself
                          passes its own discharge; no effect with local time stepping, as in the reference */
} fs_params;

/* clock fields of GPUSolver: t, steps, lastDt, dtAvg (solver-gpu.js:578-586) */
typedef struct fs_clock {
  double t;
  double lastDt;
  double dtAvg;
  int64_t steps;
  int64_t cellUpdates; /* interior open cells advanced, summed over steps (LTS: per own-level step) */
} fs_clock;

/* readState().mass of GPUSolver (solver-gpu.js:683-689), FP64 books */
typedef struct fs_mass {
  double volume, rain, infil, outflow, exchanged, clipped, expected, rel, relExClip, initial;
  double giStored;     /* water the GI stores hold, m^3 (counted in infil, as the reference books it) */
} fs_mass;

typedef struct fs_pool *fs_pool_t;

/* ---- one solver: new GPUSolver(device, o) (solver-gpu.js:35) ----
 * z, mask required (mask may be NULL = all open); manning NULL uses manningScalar;
 * infil (m/s) and rainField NULL = none / 1; inlets are interior indices j*nx+i. */
typedef struct fs_solver *fs_handle;
int fs_create(const fs_params *p, const double *z, const double *manning, double manningScalar,
              const uint8_t *mask, const double *infil, const double *rainField, const int32_t *inlets,
              int32_t nInlets, fs_handle *out);
void fs_destroy(fs_handle s);                              /* destroy() (solver-gpu.js:692) */
const char *fs_last_error(fs_handle s);                    /* NULL handle: last create error */
/* setDepth + setVelocity (solver-gpu.js:301-318): host-evaluated h, hu, hv; bookInitial=1 books
 * vol.initial = sum of the stored FP32 depths over open cells, as setDepth does */
int fs_set_state(fs_handle s, const double *h, const double *hu, const double *hv, int32_t bookInitial);
int fs_set_rain(fs_handle s, double rate_ms);              /* rainRate property (solver-gpu.js:48, 603) */
/* setRainField (solver-gpu.js:470-478); a solver created without a field gets a plane of its own (nx*ny*4
   bytes, within the pool's HBM budget, FS_E_BUDGET beyond it) */
int fs_set_rain_field(fs_handle s, const double *field);
int fs_set_inlet_rates(fs_handle s, const double *q, int32_t n);           /* setInletRates (:557) */
/* the stageCells option (solver-gpu.js:125-133): interior edge cells j*nx+i whose outward ghosts hold
 * the stage; a new list replaces the old one */
int fs_set_stage_cells(fs_handle s, const int32_t *cells, int32_t n);
int fs_set_stage(fs_handle s, double z);                   /* setStage (solver-gpu.js:554), m absolute */
/* the crest option (solver-gpu.js:18-20, 144-151): flood-barrier faces, axis 0 = the face between cells
 * (i-1, j) and (i, j), axis 1 = between (i, j-1) and (i, j), both interior; crest z (m absolute, non-finite
 * = not deployed), weir coefficient C (m^0.5/s; the reference's default 1.6), ground (m absolute,
 * non-finite = unknown); a new list replaces the old one and restarts the books */
int fs_set_crests(fs_handle s, const int32_t *i, const int32_t *j, const int32_t *axis, const double *z,
                  const double *C, const double *ground, int32_t n);
int fs_set_crest_levels(fs_handle s, const double *z);     /* setCrestLevels (solver-gpu.js:486-488) */
/* readBarriers (solver-gpu.js:490-504): 9 doubles per face: plus, minus (m^3), peak (m^2/s), level (m,
 * -inf = never wet), load (kN/m), tPeak, tLevel, tLoad, tOver (s, NaN = never overtopped) */
int fs_read_barriers(fs_handle s, double *out);
/* the ga option (solver-cpu.js:96): Green-Ampt soil per cell (psi dtheta m, Fmax m, pervious share; cells
 * with perv > 0); the infiltration capacity K of the infil layer becomes K (1 + psi dtheta / F). The soil and
 * GI planes (16 bytes per cell each) and the barrier crests need a grid of at most 2^28 cells (FS_E_ARG beyond) */
int fs_set_ga(fs_handle s, const double *psiD, const double *fmax, const double *perv);
/* the gi option (solver-cpu.js:95): GI stores per cell (intake m/s, capacity m, drain m/s; cells with cap > 0) */
int fs_set_gi(fs_handle s, const double *rate, const double *cap, const double *drain);
/* readInlets (:564-576): per inlet the volume exchanged since the last read (m^3) and the cell's depth
 * after its last exchange (m); both restart at zero, as the reference clears its INX record */
int fs_read_inlets(fs_handle s, double *vol, double *h, int32_t n);
/* readInlets().app (:564-576; approachAt, kernels.wgsl.js:620-637): per inlet, after its last exchange, the
 * frontal and side discharge arriving at its cell (m^3/s), the depth of the frontal neighbour (m) and its
 * axis (0 x, 1 y); app holds 4*n floats */
int fs_read_inlet_approach(fs_handle s, float *app, int32_t n);
int fs_advance(fs_handle s, double tTarget, int32_t K, fs_clock *clk);     /* advance (:589-604) */
/* fs_advance without any allocation (no counterpart in the reference): for callers that cannot allocate
 * device memory, such as tight serving loops and harnesses that forbid hipMalloc. The same kernels in the
 * same order as fs_advance, launched plainly (no captured graphs), with the solver's descriptor, work list
 * and per-wavefront partials in a fixed 4 MiB device array that the library holds on each device until the
 * process ends (a solver of up to ~52,000 work items, ~20,000 with local time stepping: any grid up to
 * 8192 x 8192 at the default 64 rows per strip, with or without it); a larger grid returns FS_E_ARG. Results are bit for bit those
 * of fs_advance. Every call on one device shares that array: calls that overlap on one device (other
 * threads, other streams) race, so they must be serialised by the caller. Timing mode (fs_pool_set_timing)
 * must be off (FS_E_STATE), and the solver's device must be current. */
int fs_advance_noalloc(fs_handle s, double tTarget, int32_t K, fs_clock *clk);
/* frame (:607-622): at most K steps toward tEnd (K = 0: none), then the readout (pack, kernels.wgsl.js:
 * 1346-1355): out (nx*ny*2 floats: depth, speed) may be NULL */
int fs_frame(fs_handle s, double tEnd, int32_t K, float *out, fs_clock *clk);
/* readState (:651-690): interior FP64 copies (any may be NULL) and the FP64 books */
int fs_read_state(fs_handle s, double *h, double *hu, double *hv, double *hmax, fs_mass *mass);
int fs_clock_get(fs_handle s, fs_clock *clk);
/* LTS tile levels of the last cycle (8x8 tiles, tx*ty bytes, row-major): the reference keeps them
 * in its TL buffer (kernels.wgsl.js:307) and CPUSolver.LV (solver-cpu.js:124) */
int fs_lts_levels(fs_handle s, uint8_t *levels, int32_t *tx, int32_t *ty);
/* readHazard (solver-gpu.js:515-529): per interior cell peak d*v (m^2/s), peak force per width (kN/m),
 * time deeper than 15 cm (s), its first time (s, NaN never), peak speed (m/s), first time deeper than
 * 5 cm (s, NaN never), worst ARR 2019 H1-H6, FD2320 and FD2320-urban classes (0 never classed); any
 * output may be NULL. Needs fs_params.hazard = 1. */
int fs_read_hazard(fs_handle s, float *dv, float *force, float *duration, float *arrival, float *vmax, float *tWet,
                   uint8_t *arr, uint8_t *defra, uint8_t *defraUrban);
/* readInfiltration (solver-gpu.js:507-512): depth infiltrated so far per interior cell under Green-Ampt
 * (m per pervious area; 0 without it), nx*ny floats */
int fs_read_infiltration(fs_handle s, float *F);
/* tileStats (:625-629): the active-tile counts of the last step. Global stepping skips quiet dry strips
 * (58 columns x rows cells, the unit of this build's active tiles): tiles = strips, update = flux = strips the
 * last step advanced (a strip evaluates its faces in the pass that updates it, so the reference's two lists,
 * update and flux, are one set here). tiled = 0 when every strip steps every step (tiles: false, GI stores).
 * Local time stepping: tiles = the 8x8 level tiles, every one of which steps each cycle (update = flux =
 * tiles). cellsActive: open cells the last step advanced; cellsTotal: nx*ny. */
typedef struct fs_tiles {
  int32_t tiled, pad;
  int64_t tiles, update, flux, cellsActive, cellsTotal;
} fs_tiles;
int fs_tile_stats(fs_handle s, fs_tiles *out);

/* ---- batched / served instances (no counterpart in the reference, which runs one
 * solve at a time: src/batch.js:1-4). Many uneven-size instances share one launch
 * sequence; admission is against an HBM budget. ---- */
int fs_pool_create(size_t hbmBudgetBytes, fs_pool_t *out);
void fs_pool_destroy(fs_pool_t pool);
const char *fs_pool_last_error(fs_pool_t pool);
size_t fs_pool_bytes_needed(const fs_params *p, int32_t withInfil, int32_t withRainField);
int fs_pool_add(fs_pool_t pool, const fs_params *p, const double *z, const double *manning, double manningScalar,
                const uint8_t *mask, const double *infil, const double *rainField, const int32_t *inlets,
                int32_t nInlets, int32_t *slot);
int fs_pool_remove(fs_pool_t pool, int32_t slot);
/* rows per wavefront strip of a global-stepping instance (8..4096), the tile shape the autotuner
 * picks (floodsim_amd/batch.py); results do not depend on it */
int fs_pool_set_rows(fs_pool_t pool, int32_t slot, int32_t rows);
int fs_pool_set_state(fs_pool_t pool, int32_t slot, const double *h, const double *hu, const double *hv,
                      int32_t bookInitial);
int fs_pool_set_rain(fs_pool_t pool, int32_t slot, double rate_ms);
int fs_pool_set_inlet_rates(fs_pool_t pool, int32_t slot, const double *q, int32_t n);
int fs_pool_set_stage_cells(fs_pool_t pool, int32_t slot, const int32_t *cells, int32_t n);
int fs_pool_set_stage(fs_pool_t pool, int32_t slot, double z);
int fs_pool_set_crests(fs_pool_t pool, int32_t slot, const int32_t *i, const int32_t *j, const int32_t *axis,
                       const double *z, const double *C, const double *ground, int32_t n);
int fs_pool_set_crest_levels(fs_pool_t pool, int32_t slot, const double *z);
int fs_pool_read_barriers(fs_pool_t pool, int32_t slot, double *out);
int fs_pool_set_ga(fs_pool_t pool, int32_t slot, const double *psiD, const double *fmax, const double *perv);
int fs_pool_set_gi(fs_pool_t pool, int32_t slot, const double *rate, const double *cap, const double *drain);
/* advance the listed slots together until each reaches its own target (K steps per launch group);
 * clks (n entries) may be NULL */
int fs_pool_advance(fs_pool_t pool, const int32_t *slots, const double *tTargets, int32_t n, int32_t K,
                    fs_clock *clks);
/* exactly K steps of every listed slot (stopping at its target), one device sync */
int fs_pool_step(fs_pool_t pool, const int32_t *slots, const double *tTargets, int32_t n, int32_t K);
int fs_pool_read_state(fs_pool_t pool, int32_t slot, double *h, double *hu, double *hv, double *hmax,
                       fs_mass *mass);
int fs_pool_clock(fs_pool_t pool, int32_t slot, fs_clock *clk);
size_t fs_pool_bytes_used(fs_pool_t pool);
/* wall time in ms of the dominant (second-stage) kernel, averaged over the launches since the
 * last call, measured with HIP events on the pool's stream; -1 when not measured */
int fs_pool_kernel_ms(fs_pool_t pool, double *stage1_ms, double *stage2_ms, int64_t *launches);
int fs_pool_set_timing(fs_pool_t pool, int32_t on);
/* timing mode: total ms of the dominant kernel's launches (the second RK stage with the sources;
 * under LTS one launch per record), the open cells those launches advanced, and the launch count,
 * since the last call (no counterpart in the reference: the roofline measurement) */
int fs_pool_dominant_kernel(fs_pool_t pool, double *ms_total, double *cells_total, int64_t *launches);
/* Test switch (no counterpart in the reference): FS_DEBUG_NO_STEPS = 1 makes every launch group of the pool
 * run no steps, so the clock never moves and an advance ends with FS_E_STALL and the reference's
 * no-progress message (solver-gpu.js:601) after its progress checks, with no kernel left running; 0 restores
 * normal stepping */
#define FS_DEBUG_NO_STEPS 1
int fs_pool_debug(fs_pool_t pool, int32_t what, int32_t value);
/* the underlying pool of a single solver (for timing and tuning calls) */
fs_pool_t fs_solver_pool(fs_handle s);

/* HBM reference: copies `bytes` device-to-device `iters` times with a plain vector copy kernel
 * and reports read+write GB/s and ms per copy (no counterpart in the reference: it measures the
 * achievable bandwidth the roofline is quoted against, and calibrates byte counters) */
int fs_copy_probe(size_t bytes, int32_t iters, double *gbps, double *ms);

/* Device-state snapshot / restore: a host copy of everything the solver holds on the device (state,
 * layers, books, LTS stores, hazard products, barrier crests and books, Green-Ampt and GI planes, a rain
 * field given after creation) plus its clock and FP64 books, restorable into the same solver or one built
 * from the same model and options (gpuSnapshot / gpuRestore / solverState, checkpoint.js:94-113, save every
 * storage buffer of the solver and its JS state; a restore into other buffer sizes is an error there too).
 * The buffer is opaque, fs_snapshot_bytes long. */
size_t fs_snapshot_bytes(fs_handle s);
int fs_snapshot(fs_handle s, void *out, size_t bytes);
int fs_restore(fs_handle s, const void *in, size_t bytes);
/* The parts of that buffer, in order, as the named buffers gpuSnapshot returns and gpuRestore size-checks
 * (checkpoint.js:96-112): bytes[FS_SNAP_RECORD] the clock, step record and FP64 books; bytes[FS_SNAP_PLANES]
 * the solver's main device allocation (state, layers, LTS stores, hazard products, strip tables);
 * bytes[FS_SNAP_BARRIERS] the barrier crests and books; bytes[FS_SNAP_SOIL] the Green-Ampt and GI planes;
 * bytes[FS_SNAP_RAINFIELD] a rain field given after creation. A part the solver does not hold is 0 bytes.
 * The parts add up to fs_snapshot_bytes. */
enum { FS_SNAP_RECORD = 0, FS_SNAP_PLANES, FS_SNAP_BARRIERS, FS_SNAP_SOIL, FS_SNAP_RAINFIELD, FS_SNAP_PARTS };
int fs_snapshot_layout(fs_handle s, uint64_t bytes[FS_SNAP_PARTS]);
size_t fs_pool_snapshot_bytes(fs_pool_t pool, int32_t slot);
int fs_pool_snapshot(fs_pool_t pool, int32_t slot, void *out, size_t bytes);
int fs_pool_restore(fs_pool_t pool, int32_t slot, const void *in, size_t bytes);

/* library version string */
const char *fs_version(void);

#ifdef __cplusplus
}
#endif
#endif

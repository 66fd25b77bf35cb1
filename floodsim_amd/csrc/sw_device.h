// Device-side data of the shallow-water step: instance descriptors, step records, partials.
// Shared by the kernels (sw_kernels.hip) and the host runtime (floodsim.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fs {

constexpr int kLanes = 64;                   // wavefront width on CDNA
constexpr int kHalo = 3;                     // theta limiter + MUSCL need 3 cells each side
constexpr int kOwn = kLanes - 2 * kHalo;     // columns a wavefront strip owns (58)
constexpr int kLtsTile = 8;                  // LTS level tiles (solver-cpu.js:14)
constexpr int kMaxLevels = 4;                // LTS levels 1..4 (solver-cpu.js:121)
constexpr double kLtsSafe = 0.9;             // share of its CFL step a tile may take (solver-cpu.js:15)
constexpr int kSpeedSlots = 4096;            // LTS: slots of the per-block tile-speed maxima
// wavefronts per block of the strip kernels (global stepping) and of the LTS record kernels: a block holds
// its CU slots until its longest wavefront ends, so blocks of strips or parts of uneven length keep fewer
// waves resident the more wavefronts they hold. One per block, measured at 8192^2: LTS cycle 20.43 ms with 4,
// 19.26 with 2, 19.01 with 1; global step 2.085 / 2.102 / 2.058 ms.
#ifndef FS_WPB_G
#define FS_WPB_G 1
#endif
#ifndef FS_WPB_L
#define FS_WPB_L 1
#endif
constexpr int kGlobalWaves = FS_WPB_G;
constexpr int kLtsWaves = FS_WPB_L;

enum { BC_WALL = 0, BC_OPEN = 1, BC_OUTFLOW = 2 };

// Device loads and stores through the global address space from a wave-uniform row base plus a
// 32-bit lane offset: global_load/global_store with an SGPR base and a 32-bit VGPR offset, no
// 64-bit per-lane address arithmetic. A plane holds at most 2^32 bytes (checked at creation).
#define FS_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ T gld(const T *row, int x) {
  return *(const FS_GLOBAL T *)((const FS_GLOBAL char *)row + (uint32_t)x * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void gst(T *row, int x, T v) {
  *(FS_GLOBAL T *)((FS_GLOBAL char *)row + (uint32_t)x * (uint32_t)sizeof(T)) = v;
}
// HIP's vector types are classes bound to the generic address space: move them as clang vectors
typedef float fs_v2 __attribute__((ext_vector_type(2)));
typedef float fs_v4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 gld(const float4 *row, int x) {
  const fs_v4 v = *(const FS_GLOBAL fs_v4 *)((const FS_GLOBAL char *)row + (uint32_t)x * 16u);
  return make_float4(v.x, v.y, v.z, v.w);
}
typedef int fs_i2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ int2 gld(const int2 *row, int x) {
  const fs_i2 v = *(const FS_GLOBAL fs_i2 *)((const FS_GLOBAL char *)row + (uint32_t)x * 8u);
  return make_int2(v.x, v.y);
}
// the first two components of a 16-byte element (an 8-byte load)
__device__ __forceinline__ float2 gld_lo(const float4 *row, int x) {
  const fs_v2 v = *(const FS_GLOBAL fs_v2 *)((const FS_GLOBAL char *)row + (uint32_t)x * 16u);
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ void gst(float4 *row, int x, float4 v) {
  const fs_v4 t = {v.x, v.y, v.z, v.w};
  *(FS_GLOBAL fs_v4 *)((FS_GLOBAL char *)row + (uint32_t)x * 16u) = t;
}
__device__ __forceinline__ void gst(float2 *row, int x, float2 v) {
  const fs_v2 t = {v.x, v.y};
  *(FS_GLOBAL fs_v2 *)((FS_GLOBAL char *)row + (uint32_t)x * 8u) = t;
}
// mask byte bits: open (not a wall) cell, inlet cell, Manning n > 0 (the step-flux rules apply), stage
// cell (its outward ghosts hold the stage)
enum { MASK_OPEN = 1, MASK_INLET = 2, MASK_ROUGH = 4, MASK_STAGE = 8 };
// DevInst flags: infiltration layer, rain field, implicit friction and step flux (F26), stage cells
enum { FLAG_INFIL = 1, FLAG_RF = 2, FLAG_IMPF = 4, FLAG_STEPF = 8, FLAG_STAGE = 16 };

// Per-instance mutable record: clock, dt, FP64 books. Written by dt_final only (and the host
// between launches); the stage kernels read dt and the rain rate from it.
struct StepRec {
  double t;        // time at the start of the pending step
  double tAfter;   // time once the pending step is done (exactly tEnd when clamped)
  double tEnd;     // target of the current advance
  double rain;     // FP64 books, m^3
  double infil;
  double outflow;
  double clipped;
  double exchanged;
  double rainArea; // m^2: sum of the rain multiplier over open cells times dx^2
  double cellUpdates;
  float dt;        // pending step (0 = no-op)
  float smax;      // last max wave speed
  float lastDt;
  float rainRate;  // m/s
  int64_t steps;
  int32_t rainCells;  // open cells with a uniform multiplier of 1 (no rain field), else -1
  int32_t openCells;
  int32_t levels;     // LTS levels (0 = global)
  int32_t dirty;      // the host changed the state: the next dt comes from a fresh speed reduction
  // local time stepping: the current cycle (one "step" of an LTS instance is one cycle)
  double cycT0;       // cycle start
  double cycD;        // cycle length
  double base;        // base step delta = D / 2^L
  double lrf[kMaxLevels + 1];     // sum of rain multipliers over own open cells per level
  float ldt[kMaxLevels + 1];      // record step per level: delta * 2^m (0 = no-op)
  int32_t lcells[kMaxLevels + 1]; // own open cells per level
  int32_t lcount[kMaxLevels + 1]; // parts per level (the records' work lists)
  int32_t records;                // records run so far (diagnostic)
  int32_t active;                 // open cells the last step advanced (dry strips skipped)
  int32_t activeStrips;           // strips the last step advanced (dry-strip skipping on)
  float stage;                    // water level of the stage cells' ghosts, m relative to the bed minimum
};

// one inlet: cell index (y*pitch + x), rate (m^3/s, + into the network, - a point inflow), the cell's depth
// after its last exchange, and the volume exchanged since the last read (the reference's INX record,
// kernels.wgsl.js:597-618, cleared by readInlets, solver-gpu.js:564-576)
struct Inlet {
  int cell;
  float q;
  float h;      // depth after the last exchange, m
  float pad;
  double vol;   // exchanged since the last read, m^3
};

// per wavefront strip and stage: boundary outflow (m^3), clipped (m^3), infiltration (m^3), max
// speed, open cells advanced (second stage)
struct Part {
  double out, clip, infil;
  float smax;
  int32_t cells;
};

struct DevInst {
  float *h, *hu, *hv, *w;       // A: U^n, and U^{n+1} after the second stage (in place)
  float *h1, *hu1, *hv1, *w1;   // B: U* (first-stage result)
  int ab;                       // cells from a variable's A plane to its B plane (h1 = h + ab, hu1 = hu + ab, ...)
  const float *z;               // bed, m, relative to the instance minimum
  const float *n;               // Manning n
  const uint8_t *mask;          // MASK_* bits
  float *hmax;
  const float *infil;           // m/s, or null
  const float *rf;              // rain multiplier, or null
  StepRec *rec;
  int nx, ny, pitch, flags;
  float dx, inv_dx, g, hdry, hfront, cfl, dtDry, area;
  int secondOrder, limiter;
  int bc[4];
  int wbeg, wend;               // this instance's range in the work list
  Inlet *inl;                   // this instance's inlets
  float *ispeed;                // wave speed of each inlet cell after the exchange
  float4 *iapp;                 // water arriving at each inlet cell after its exchange: frontal and side
                                // discharge (m^3/s), depth of the frontal neighbour (m), its axis (0 x, 1 y)
  int nin;
  int hl, own;                  // strip geometry: halo lanes on the left, owned columns
  int rows, sxn;                // rows per strip, strips across
  // dry-strip skipping (global stepping; null = every strip every step). A strip is quiet when
  // all its cells hold h = w = hu = hv = 0; a step advances a strip only if a strip of its 3x3
  // block is not quiet or rain falls on it, which leaves the results identical to advancing all.
  uint8_t *wet;                 // per strip: not quiet after the last step (+ inlet exchange)
  uint8_t *act;                 // per strip: advanced in the current step
  const uint8_t *srain;         // per strip: bit 0 an open cell with a positive rain multiplier, bit 1 a stage cell
  const int32_t *sopen;         // per strip: open cells it owns
  // hazard products (null: not tracked): A = (peak d*v, peak force, time deeper than 15 cm, first
  // time deeper than 15 cm or -1), B = (peak speed, first time deeper than VCUT or -1, worst classes
  // packed ARR + 8 FD2320 + 64 FD2320-urban, 0)
  float4 *hzA, *hzB;
  // local time stepping (L > 0)
  int L, tx, ty;                // levels; LTS tiles across and down (8x8 cells)
  uint8_t *lv, *lv0;            // tile levels (smoothed, raw)
  float *tsp;                   // tile max wave speed at the cycle start
  unsigned *speedSlots;         // kSpeedSlots block maxima of the tile speeds (float bits), zero between cycles
  const int32_t *tileOpen;      // open cells per tile
  const float *tileRf;          // rain multiplier summed over the open cells of a tile (rain field only)
  float4 *st4;                  // flux a waiting cell has gathered: (h, its Kahan residual, hu, hv), one 16-byte load
  float4 *g4;                   // a waiting cell's record: its first-stage gather (h, hu, hv; added to the store
                                // after) and its store with that gather, the second stage's theta reads (second order)
  uint8_t *gflag;               // first order: 1 = the cell gathered in the current record
  float4 *bsx, *bsy;            // a cell's reconstruction toward its faces at their last evaluation:
                                // x (h west, dz west, h east, dz east), y (south, north); dz = face bed - cell bed
  uint64_t *urows;              // per level and work unit (stride wend - wbeg): the unit's rows a record of the
                                // level owns (bit q: row q of the unit), gaps of up to 2 kHalo rows filled
  int32_t *ucells;              // per work unit and level: own open cells
  int4 *plist;                  // per level (stride capParts): the parts of the level's records, (0, first
                                // column, first row, rows) as a work item
  int capParts;                 // parts a level's list holds (8 per unit)
  double *rend;                 // a record's books per slice of its partials (3 x the slices)
  unsigned *ticket;             // slices summed so far in the current record (0 between records)
  // flood barriers (null: none). Per cell: crest and weir coefficient of its west face (x) and south face
  // (y), crest relative to the bed minimum (-1e30 = none); the barrier index of those faces (-1 = none).
  // Per barrier: (cell k of its right/upper side, k of the other side, axis 0 = x face / 1 = y face, 0),
  // the ground under it (-1e30 = unknown), the theta-scaled mass flux of the last stage, and its books
  // (plus, minus m^3; peak m^2/s; level m; load kN/m; their times and the first overtopping, s)
  float4 *crest;
  int2 *bidx;
  const int4 *barFace;
  const float *barGround;
  float *barFlux;
  double *barBook;
  int nbar;
  // Green-Ampt soil (psi dtheta, Fmax, pervious share, F infiltrated so far) and GI stores (intake rate,
  // capacity, drain rate, stored depth) per cell; null: none
  float4 *soil, *gi;
};



}  // namespace fs

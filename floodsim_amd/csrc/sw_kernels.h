// HIP kernels of the shallow-water step for gfx950 (CDNA4).
//
// One RK stage is one kernel. Each 64-lane wavefront owns a strip of 58 columns (lanes 3..60;
// 3 halo lanes each side) and walks its rows top to bottom with a register window:
//   row r loaded -> x reconstruction and x faces of row r (x neighbours by DPP lane shifts)
//                -> y reconstruction of row r-1, y face (r-2 | r-1)
//                -> draining-time limiter theta of row r-2
//                -> conservative update (and, in the second stage, sources) of row r-3
// so every face, theta and update is computed from values held in registers: nothing but
// the state, the static layers and the outputs crosses HBM, and no LDS or barrier is used.
//
// Physics follow the reference GPU kernels (src/kernels.wgsl.js) and CPU solver
// (src/solver-cpu.js) cited per function. Elevations enter only as differences of
// neighbouring FP32 beds (exact by Sterbenz for neighbours within a factor of two) plus
// depth-relative terms, never as absolute eta = h + z, so thin films on high ground keep
// their digits.
#pragma once
#include <type_traits>
#include "sw_device.h"

// which plain stage kernels issue a row pass's loads together (bit 0: global second stage, bit 1: LTS first
// stage, bit 2: LTS second stage, bit 3: global first stage)
#ifndef FS_PRE
#define FS_PRE 11
#endif

namespace fs {

// ---------------------------------------------------------------- lane shifts (DPP)
// (bound_ctrl: the lane without a source, 0 or 63, reads 0, a halo lane whose results are unused;
// the shift can then fold into the instruction that consumes it)
__device__ __forceinline__ float from_left(float v) {   // lane l receives lane l-1
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float from_right(float v) {  // lane l receives lane l+1
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ int from_left_i(int v) {
  return __builtin_amdgcn_mov_dpp(v, 0x138, 0xf, 0xf, true);
}
__device__ __forceinline__ int from_right_i(int v) {
  return __builtin_amdgcn_mov_dpp(v, 0x130, 0xf, 0xf, true);
}

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsq(float x) { return __builtin_amdgcn_sqrtf(x); }
// x^p for x > 0 through v_log_f32 / v_exp_f32 (base 2)
__device__ __forceinline__ float fpow(float x, float p) {
  return __builtin_amdgcn_exp2f(p * __builtin_amdgcn_logf(x));
}

// minmod (solver-cpu.js:17-20; kernels.wgsl.js:83-84): of two slopes of one sign the one nearer zero, else 0,
// which is the median of (a, b, 0): one v_med3_f32. (Where a * b would underflow FP32 the product test of the
// reference's form gives 0 and the median the smaller slope, as the FP64 product does.)
__device__ __forceinline__ float minmod(float a, float b) { return __builtin_amdgcn_fmed3f(a, b, 0.f); }

// slope limiter: minmod or van Leer (solver-cpu.js:17-24; kernels.wgsl.js:83-87)
__device__ __forceinline__ float limit(float a, float b, int vanleer) {
  const float p = a * b;
  const float v = p <= 0.f ? 0.f : 2.f * p * frcp(a + b);
  return vanleer ? v : minmod(a, b);
}

// HLLC in the face-normal frame with dry-bed speed estimates (solver-cpu.js:579-602;
// kernels.wgsl.js:183-210). Branch-free: every path is evaluated and selected.
__device__ __forceinline__ void hllc(float hL, float uL, float vL, float hR, float uR, float vR, float g,
                                     float hdry, float &F0, float &F1, float &F2) {
  const bool wetL = hL > hdry, wetR = hR > hdry;
  uL = wetL ? uL : 0.f; vL = wetL ? vL : 0.f;
  uR = wetR ? uR : 0.f; vR = wetR ? vR : 0.f;
  const float cL = fsq(g * hL), cR = fsq(g * hR);
  const float us = 0.5f * (uL + uR) + cL - cR;
  const float cs = 0.5f * (cL + cR) + 0.25f * (uL - uR);
  float SL = fminf(uL - cL, us - cs), SR = fmaxf(uR + cR, us + cs);
  SL = wetR ? SL : uL - cL;
  SR = wetR ? SR : uL + 2.f * cL;
  SL = wetL ? SL : uR - 2.f * cR;
  SR = wetL ? SR : uR + cR;
  const float qL = hL * uL, qR = hR * uR;
  const float FL1 = qL * uL + 0.5f * g * hL * hL;
  const float FR1 = qR * uR + 0.5f * g * hR * hR;
  const float inv = frcp(SR - SL);
  const float H0 = (SR * qL - SL * qR + SL * SR * (hR - hL)) * inv;
  const float H1 = (SR * FL1 - SL * FR1 + SL * SR * (qR - qL)) * inv;
  // contact speed SM >= 0 <=> numerator and denominator share a sign (no division needed)
  const float dSM = hR * (uR - SR) - hL * (uL - SL);
  const float nSM = SL * hR * (uR - SR) - SR * hL * (uL - SL);
  const bool left = !(fabsf(dSM) > 1e-30f) || (nSM * dSM >= 0.f);
  const float H2 = H0 * (left ? vL : vR);
  float f0 = H0, f1 = H1, f2 = H2;
  if (SL >= 0.f) { f0 = qL; f1 = FL1; f2 = qL * vL; }
  if (SR <= 0.f && !(SL >= 0.f)) { f0 = qR; f1 = FR1; f2 = qR * vR; }
  const bool dry = !wetL && !wetR;
  F0 = dry ? 0.f : f0;
  F1 = dry ? 0.f : f1;
  F2 = dry ? 0.f : f2;
}

// Face with hydrostatic reconstruction (Audusse) and wall mirroring (solver-cpu.js:237-267;
// kernels.wgsl.js:233-263). L and R are the reconstructed (h, eta - z_cell, un, ut) of the two
// cells; dz = z_R - z_L of the cell beds. The face beds are z_cell + (e - h); only their
// difference delta enters: h*_L = max(0, hL - max(0, delta)), h*_R = max(0, hR - max(0, -delta)).
struct Face { float F0, F1, F2, cL, cR, hLs, hRs; };

// EXT (extended faces), for instances with the step flux or flood barriers:
// - a barrier crest on the face (dcL = crest - z_L cell bed; -1e30 = none) joins the face beds in the
//   hydrostatic reconstruction, z* = max(zL, zR, crest) (solver-cpu.js:244-246; kernels.wgsl.js:240-242);
// - the step-flux rule (F26, solver-cpu.js:250-273; kernels.wgsl.js:247-257) for faces between two open
//   cells with friction (RL, RR: Manning n > 0; both false on a weir face): a wet film whose reconstructed
//   depth over the other side's higher face bed is zero passes its own discharge upwind, and the lower cell
//   feels the whole step
template <bool EXT = false>
__device__ __forceinline__ Face face(float hL, float eL, float uL, float vL, bool ML, float hR, float eR, float uR,
                                     float vR, bool MR, float dz, float g, float hdry, bool RL = false,
                                     bool RR = false, float dcL = -1e30f) {
  const float dcRel = ML ? dcL : dcL - dz;   // crest over the bed of the cell the left state comes from
  if (!MR) { hR = hL; eR = eL; uR = -uL; vR = vL; }
  if (!ML) { hL = hR; eL = eR; uL = -uR; vL = vR; }
  const float delta = (ML && MR) ? (dz + (eR - hR)) - (eL - hL) : 0.f;
  // z* above the left and right face beds
  float sL = fmaxf(0.f, delta), sR = fmaxf(0.f, -delta);
  if constexpr (EXT) {
    const float cOver = dcRel - (eL - hL);   // crest above the left face bed
    sL = fmaxf(sL, cOver);
    sR = fmaxf(sR, cOver - delta);
  }
  const float hLs = fmaxf(0.f, hL - sL);
  const float hRs = fmaxf(0.f, hR - sR);
  Face f;
  hllc(hLs, uL, vL, hRs, uR, vR, g, hdry, f.F0, f.F1, f.F2);
  const float hg = 0.5f * g;
  f.cL = ML ? hg * (hL * hL - hLs * hLs) : 0.f;
  f.cR = MR ? hg * (hR * hR - hRs * hRs) : 0.f;
  f.hLs = hLs;
  f.hRs = hRs;
  if constexpr (EXT) {
    if (ML && MR && RL && RR) {
      // the surface of the lower side lies z* - (its face bed) - h under the step top
      if (hLs > hdry && hRs <= 0.f && hR > hdry) {
        const float u = fmaxf(uL, 0.f), q = hLs * u;
        f.F0 = q; f.F1 = q * u + hg * hLs * hLs; f.F2 = q * vL;
        f.cR = hg * hR * hR + g * hR * (sR - hR);
      } else if (hRs > hdry && hLs <= 0.f && hL > hdry) {
        const float u = fminf(uR, 0.f), q = hRs * u;
        f.F0 = q; f.F1 = q * u + hg * hRs * hRs; f.F2 = q * vR;
        f.cL = hg * hL * hL + g * hL * (sL - hL);
      }
    }
  }
  if (!ML && !MR) { f.F0 = f.F1 = f.F2 = f.cL = f.cR = 0.f; }
  return f;
}

// Flow over a barrier crest zc whose crest stands above both cell beds (the caller checks): broad-crested
// weir q = C H1^1.5 with Villemonte submergence (1 - (H2/H1)^1.5)^0.385, capped so one step cannot more than
// half-close the level difference; the momentum carries the critical velocity over the crest, the pressure
// is the mean of the two sides' depths over the crest (weir, kernels.wgsl.js:212-230; _weir,
// solver-cpu.js:277-287). Cell-centred surfaces eA, eB (relative to the bed minimum), depths, tangential
// velocities; overwrites the face's HLLC flux.
__device__ __forceinline__ void weir_flux(Face &f, float eA, float hA, float tA, float eB, float hB, float tB, float zc,
                                          float cw, float g, float hdry, float dx, float dt) {
  const bool left = eA >= eB;
  const float eU = left ? eA : eB, hU = left ? hA : hB, tU = left ? tA : tB, eD = left ? eB : eA;
  const float pres = 0.25f * g * (f.hLs * f.hLs + f.hRs * f.hRs);
  const float H1 = eU - zc;
  f.F0 = 0.f; f.F1 = pres; f.F2 = 0.f;
  if (H1 <= hdry || hU <= hdry) return;
  const float r = fmaxf(0.f, eD - zc) / H1;
  if (r >= 1.f) return;
  float q = cw * fpow(H1, 1.5f) * fpow(1.f - fpow(fmaxf(r, 1e-30f), 1.5f), 0.385f);
  if (r <= 0.f) q = cw * fpow(H1, 1.5f);
  if (dt > 0.f) q = fminf(q, (eU - eD) * dx / (4.f * dt));
  const float un = q / fmaxf(2.f / 3.f * H1, hdry);
  const float F0 = left ? q : -q;
  f.F0 = F0; f.F1 = q * un + pres; f.F2 = F0 * tU;
}

// a product and a sum rounded on their own (never fused into a multiply-add with their neighbours): values
// that earlier builds stored to memory and read back keep their bits
__device__ __forceinline__ float mul_nc(float a, float b) {
#pragma clang fp contract(off)
  return a * b;
}
__device__ __forceinline__ float add_nc(float a, float b) {
#pragma clang fp contract(off)
  return a + b;
}

// Kahan-compensated depth: stored h with residual w, true depth h - w (kernels.wgsl.js:265-272)
__device__ __forceinline__ void kadd(float &h, float &w, float d) {
  const float y = d - w;
  const float t = h + y;
  w = (t - h) - y;
  h = t;
}

// free-outfall edge velocity: Manning on the surface slope toward the edge beyond 1e-4, at most
// critical (solver-cpu.js:179-183; kernels.wgsl.js:92-101). e = edge cell, i = next cell inward.
__device__ __forceinline__ float out_vel(float he, float ze, float ne, float hi, float zi, bool Mi, float inv_dx,
                                         float g, float hdry) {
  if (he <= hdry || hi <= hdry || !Mi) return 0.f;
  const float s = ((hi - he) + (zi - ze)) * inv_dx - 1e-4f;
  if (s <= 0.f) return 0.f;
  const float crit = fsq(g * he);
  if (ne <= 0.f) return crit;
  return fminf(crit, fpow(he, 2.f / 3.f) * fsq(s) * frcp(ne));
}

// Primitive of cell (X, Y) of the evaluated state, ghosts included (solver-cpu.js:165-213;
// kernels.wgsl.js:122-156): walls mirror across the edge and flip the normal velocity, open
// and outflow edges copy the edge cell, outflow sets the outward Manning velocity. Ghost beds
// are the nearest interior bed (zero gradient), ghost cells are open. Corner ghosts never
// enter an interior cell's stencil and only need to be finite.
struct Prim { float h, u, v, z, hu, hv; bool M, in, R; int lv; };   // R: Manning n > 0

// F26 (implicitFriction): the second stage's fluxes see the predictor's velocity after friction implicit
// in the new speed over the step, k = (1 + sqrt(1 + 4 a |u|)) / 2 (solver-cpu.js:419-423, 443-452;
// kernels.wgsl.js:139-143); the update itself keeps U*
__device__ __forceinline__ void pre_friction(const DevInst &I, int k, float h, float dt, float &u, float &v) {
  const float nn = gld(I.n, k);
  if (h > I.hdry && nn > 0.f) {
    const float a = dt * I.g * nn * nn * frcp(fpow(h, 4.f / 3.f));
    const float ik = frcp(0.5f * (1.f + fsq(1.f + 4.f * a * fsq(u * u + v * v))));
    u *= ik;
    v *= ik;
  }
}

// level of the LTS tile holding interior cell (x, y)
__device__ __forceinline__ int tile_level(const DevInst &I, int x, int y) {
  return I.lv[(y >> 3) * I.tx + (x >> 3)];
}

// state source of cell k: the stage-1 stencil of a local-time-stepping record reads U* (B) for the
// cells of the stepping level and the frozen state (A) for every other cell
template <int STAGE, bool LTS>
__device__ __forceinline__ void cell_state(const DevInst &I, int x, int y, int m, float &h, float &hu, float &hv) {
  const size_t k = (size_t)y * I.pitch + x;
  bool useB = STAGE == 1;
  if constexpr (LTS) useB = STAGE == 1 && tile_level(I, x, y) == m;
  if (useB) { h = I.h1[k]; hu = I.hu1[k]; hv = I.hv1[k]; }
  else { h = I.h[k]; hu = I.hu[k]; hv = I.hv[k]; }
}

// EXT: the stage cells' ghosts (kept out of the plain kernels: their code in this edge path alone costs
// the plain second stage registers and spills)
template <int STAGE, bool LTS, bool EXT = false>
__device__ __forceinline__ Prim load_prim(const DevInst &I, int X, int Y, int m, float dt) {
  const int nx = I.nx, ny = I.ny;
  int sx = X, sy = Y;
  float fu = 1.f, fv = 1.f;
  const bool gx = X < 0 || X >= nx, gy = Y < 0 || Y >= ny;
  if (X < 0) { sx = I.bc[0] == BC_WALL ? -1 - X : 0; if (I.bc[0] == BC_WALL) fu = -1.f; }
  else if (X >= nx) { sx = I.bc[1] == BC_WALL ? 2 * nx - 1 - X : nx - 1; if (I.bc[1] == BC_WALL) fu = -1.f; }
  if (Y < 0) { sy = I.bc[2] == BC_WALL ? -1 - Y : 0; if (I.bc[2] == BC_WALL) fv = -1.f; }
  else if (Y >= ny) { sy = I.bc[3] == BC_WALL ? 2 * ny - 1 - Y : ny - 1; if (I.bc[3] == BC_WALL) fv = -1.f; }
  sx = min(max(sx, 0), nx - 1);
  sy = min(max(sy, 0), ny - 1);
  const int zx = min(max(X, 0), nx - 1), zy = min(max(Y, 0), ny - 1);
  const size_t k = (size_t)sy * I.pitch + sx;
  Prim p;
  float hu, hv;
  p.lv = 0;
  bool fromB = STAGE == 1;   // the state is U* (the second stage's evaluation point)
  if constexpr (LTS) {
    p.lv = tile_level(I, zx, zy);  // a ghost takes its edge cell's level (solver-cpu.js:125-128)
    // the state comes from the source cell (a mirrored ghost's may sit in another tile)
    const int ls = (gx || gy) ? tile_level(I, sx, sy) : p.lv;
    fromB = STAGE == 1 && ls == m;
    if (fromB) { p.h = I.h1[k]; hu = I.hu1[k]; hv = I.hv1[k]; }
    else { p.h = I.h[k]; hu = I.hu[k]; hv = I.hv[k]; }
  } else {
    cell_state<STAGE, false>(I, sx, sy, m, p.h, hu, hv);
  }
  p.hu = hu * fu;
  p.hv = hv * fv;
  const size_t kz = (size_t)zy * I.pitch + zx;   // ghosts take the edge cell's bed and roughness
  p.z = I.z[kz];
  p.in = !gx && !gy;
  const bool srcOpen = (I.mask[k] & MASK_OPEN) != 0;
  p.M = p.in ? srcOpen : true;
  p.R = (I.mask[kz] & MASK_ROUGH) != 0;
  const float ih = frcp(p.h);
  const bool wet = p.h > I.hdry;
  p.u = wet ? p.hu * ih : 0.f;
  p.v = wet ? p.hv * ih : 0.f;
  // a ghost copies its source cell after the source's friction
  if (STAGE == 1 && (EXT || LTS) && (I.flags & FLAG_IMPF) && fromB && srcOpen) pre_friction(I, (int)k, p.h, dt, p.u, p.v);
  if (gx != gy) {  // edge ghost (not a corner): free outfall
    int e = -1, i = -1, side = -1;
    if (X < 0 && I.bc[0] == BC_OUTFLOW) { e = 0; i = 1; side = 0; }
    else if (X >= nx && I.bc[1] == BC_OUTFLOW) { e = nx - 1; i = nx - 2; side = 1; }
    else if (Y < 0 && I.bc[2] == BC_OUTFLOW) { e = 0; i = 1; side = 2; }
    else if (Y >= ny && I.bc[3] == BC_OUTFLOW) { e = ny - 1; i = ny - 2; side = 3; }
    if (side >= 0) {
      i = max(i, 0);
      int ex, ey, ix, iy;
      if (side < 2) { ex = e; ey = zy; ix = min(i, nx - 1); iy = zy; }
      else { ex = zx; ey = e; ix = zx; iy = min(i, ny - 1); }
      const size_t ke = (size_t)ey * I.pitch + ex, ki = (size_t)iy * I.pitch + ix;
      float he, hi, t0, t1;
      cell_state<STAGE, LTS>(I, ex, ey, m, he, t0, t1);
      cell_state<STAGE, LTS>(I, ix, iy, m, hi, t0, t1);
      const float ov = out_vel(he, I.z[ke], I.n[ke], hi, I.z[ki], (I.mask[ki] & MASK_OPEN) != 0, I.inv_dx, I.g, I.hdry);
      if (side == 0) p.u = -ov;
      else if (side == 1) p.u = ov;
      else if (side == 2) p.v = -ov;
      else p.v = ov;
    }
    // water-level boundary (stageCells, setStage): the ghost of a stage cell holds the surface at the
    // stage where that stands above the ghost's bed, moving with the edge cell's velocity
    // (solver-cpu.js:197-202; kernels.wgsl.js:148-154)
    if (EXT && (I.mask[kz] & MASK_STAGE)) {
      const float hs = I.rec->stage - p.z;
      if (hs > I.hdry) {
        float he, hue, hve;
        cell_state<STAGE, LTS>(I, zx, zy, m, he, hue, hve);
        float ue = 0.f, ve = 0.f;
        if (he > I.hdry) {
          const float ie = frcp(he);
          ue = hue * ie;
          ve = hve * ie;
        }
        bool eB = STAGE == 1;
        if constexpr (LTS) eB = STAGE == 1 && tile_level(I, zx, zy) == m;
        if (STAGE == 1 && (I.flags & FLAG_IMPF) && eB && (I.mask[kz] & MASK_OPEN)) pre_friction(I, (int)kz, he, dt, ue, ve);
        p.h = hs; p.u = ue; p.v = ve; p.hu = hs * ue; p.hv = hs * ve;
      }
    }
  }
  return p;
}

// Primitive of interior cell (X, Y) for a row whose 64 lanes are all interior cells (the common
// case): row-based global loads, no ghost logic
// lvAhead: the cell's tile level, loaded by the caller a row ahead (the LTS second stage), else unused
template <int STAGE, bool LTS, bool EXT = false>
__device__ __forceinline__ Prim load_prim_row(const DevInst &I, int X, int Y, int m, float dt, int lvAhead = 0) {
  const int k = Y * I.pitch + X;
  Prim p;
  float hu, hv;
  p.lv = 0;
  if constexpr (LTS) p.lv = STAGE == 1 ? lvAhead : gld(I.lv, (Y >> 3) * I.tx + (X >> 3));
  bool useB = STAGE == 1;
  if constexpr (LTS) useB = STAGE == 1 && p.lv == m;
  if constexpr (LTS && STAGE == 1) {
    // a row mixes the stepping level's U* with the others' U^n: one load per variable at a per-lane offset
    const int ka = useB ? k + I.ab : k;
    p.h = gld(I.h, ka); hu = gld(I.hu, ka); hv = gld(I.hv, ka);
  } else if (useB) { p.h = gld(I.h1, k); hu = gld(I.hu1, k); hv = gld(I.hv1, k); }
  else { p.h = gld(I.h, k); hu = gld(I.hu, k); hv = gld(I.hv, k); }
  p.hu = hu;
  p.hv = hv;
  p.z = gld(I.z, k);
  p.in = true;
  const uint8_t mk = gld(I.mask, k);
  p.M = (mk & MASK_OPEN) != 0;
  p.R = (mk & MASK_ROUGH) != 0;
  const float ih = frcp(p.h);
  const bool wet = p.h > I.hdry;
  p.u = wet ? hu * ih : 0.f;
  p.v = wet ? hv * ih : 0.f;
  if (STAGE == 1 && (EXT || LTS) && (I.flags & FLAG_IMPF) && useB && p.M) pre_friction(I, k, p.h, dt, p.u, p.v);
  return p;
}

// MUSCL slopes of (h, eta, un, ut) from the two neighbours along one axis (solver-cpu.js:216-233;
// kernels.wgsl.js:159-180): first order when the cell or an open neighbour is thinner than hfront;
// wall neighbours are left out of the slope. dza = z_c - z_a, dzb = z_b - z_c.
struct Slopes { float sh, se, sn, st; };

// SF: a film with friction (Rc) thinner than half the bed's change across the cell is a cascade and
// stays first order (F26, solver-cpu.js:222-224; kernels.wgsl.js:166-170)
template <bool SF = false>
__device__ __forceinline__ Slopes muscl(float hc, float nc, float tc, float ha, float na, float ta, bool Ma,
                                        float dza, float hb, float nb, float tb, bool Mb, float dzb, int secondOrder,
                                        float hfront, int vanleer, bool Rc = false) {
  Slopes s = {0.f, 0.f, 0.f, 0.f};
  bool fo = !secondOrder || hc < hfront || (Ma && ha < hfront) || (Mb && hb < hfront);
  if constexpr (SF) fo = fo || (Rc && hc < 0.5f * fabsf((Mb ? dzb : 0.f) + (Ma ? dza : 0.f)));
  if constexpr (!SF) {
    // the plain variants (minmod at compile time): a first-order cell drops its left differences, and a
    // minmod with a zero operand is zero, so every slope is zero with no select of its own
    const bool ka = Ma && !fo;
    const float dhm = ka ? hc - ha : 0.f, dhp = Mb ? hb - hc : 0.f;
    s.sh = minmod(dhm, dhp);
    s.se = minmod(ka ? dhm + dza : 0.f, Mb ? dhp + dzb : 0.f);
    s.sn = minmod(ka ? nc - na : 0.f, Mb ? nb - nc : 0.f);
    s.st = minmod(ka ? tc - ta : 0.f, Mb ? tb - tc : 0.f);
    (void)vanleer;
    return s;
  }
  if (fo) return s;
  const float dhm = Ma ? hc - ha : 0.f, dhp = Mb ? hb - hc : 0.f;
  s.sh = limit(dhm, dhp, vanleer);
  s.se = limit(Ma ? dhm + dza : 0.f, Mb ? dhp + dzb : 0.f, vanleer);
  s.sn = limit(Ma ? nc - na : 0.f, Mb ? nb - nc : 0.f, vanleer);
  s.st = limit(Ma ? tc - ta : 0.f, Mb ? tb - tc : 0.f, vanleer);
  return s;
}

// wavefront reductions (fixed tree: deterministic)
__device__ __forceinline__ float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ int wave_sum_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// wavefront index with blocks of one XCD (blockIdx % 8) mapped to one contiguous run of work,
// so strips that share halo rows and columns share an L2
__device__ __forceinline__ int xcd_wave_index(int nwaves_per_block) {
  const int nb = gridDim.x, b = blockIdx.x;
  const int q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
  const int lb = x * q + min(x, r) + k;
  return lb * nwaves_per_block + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

// |u| + |v| + 2 sqrt(g h): the sum of the directional wave speeds (computeDt, solver-cpu.js:383-394)
// (contraction off: every kernel that inlines it computes the same bits, so a speed taken after the
// host writes state equals the one the second stage takes)
__device__ __forceinline__ float cell_speed(float h, float hu, float hv, float g) {
#pragma clang fp contract(off)
  const float ih = frcp(h);
  return (fabsf(hu * ih) + fabsf(hv * ih)) + 2.f * fsq(g * h);
}

// Friction, rain, infiltration and max depth of one cell after its update, then the cell's wave
// speed for the next dt (sourcesAt, kernels.wgsl.js:544-590; _sources, solver-cpu.js:456-483).
// Semi-implicit Manning: (hu, hv) /= 1 + dt g n^2 |u| / h^(4/3).
// EXT: Green-Ampt soil and GI stores (solver-cpu.js:466-474): the infiltration capacity K dt becomes
// K dt (1 + psi dtheta / F), none once the soil above the water table is full, and F grows by what went in
// over the pervious share; a GI store drains linearly, then takes what it can at its intake rate up to its
// capacity, which is booked as infiltration
// IMPF: the variant takes implicit friction (F26): the extended variants, and every LTS variant (measured:
// taking the test out of the plain LTS kernels slows their schedule)
// the per-cell layers the sources read (Manning n, max depth, mask byte, infiltration rate, rain multiplier)
struct SrcOps { float n, hmax, infil, rf; uint8_t mask; };
// SPEED: the caller takes the cell's wave speed for the next step (global stepping; local time stepping takes
// its tile speeds at the cycle's start instead), which needs the mask's inlet bit
template <bool SPEED = true>
__device__ __forceinline__ SrcOps load_src(const DevInst &I, int k) {
  SrcOps o;
  o.n = gld(I.n, k);
  o.hmax = gld(I.hmax, k);
  o.mask = SPEED ? gld(I.mask, k) : (uint8_t)0;
  o.infil = (I.flags & FLAG_INFIL) ? gld(I.infil, k) : 0.f;
  o.rf = (I.flags & FLAG_RF) ? gld(I.rf, k) : 1.f;
  return o;
}

template <bool EXT = false, bool IMPF = EXT, bool SPEED = true>
__device__ __forceinline__ void sources_cell(const DevInst &I, int k, float dt, float rain, float areaf,
                                             float &qh, float &qw, float &qhu, float &qhv, double &acc_inf,
                                             float &smax, const SrcOps *pre = nullptr) {
  const float g = I.g, hdry = I.hdry;
  const float d = qh;
  const float nn = pre ? pre->n : gld(I.n, k);
  if (d > hdry && nn > 0.f) {
    const float id = frcp(d);
    const float u = qhu * id, v = qhv * id;
    const float af = dt * g * nn * nn * frcp(fpow(d, 4.f / 3.f));
    const float sp = fsq(u * u + v * v);
    // F26 (implicitFriction): implicit in the new speed, k = (1 + sqrt(1 + 4 a |u|)) / 2 (solver-cpu.js:465-467)
    const float kf = (IMPF && (I.flags & FLAG_IMPF)) ? frcp(0.5f * (1.f + fsq(1.f + 4.f * af * sp))) : frcp(1.f + af * sp);
    qhu *= kf;
    qhv *= kf;
  }
  const float rr = (I.flags & FLAG_RF) ? rain * (pre ? pre->rf : gld(I.rf, k)) * dt : rain * dt;
  float inf = 0.f;
  if constexpr (!EXT) {
    if (I.flags & FLAG_INFIL) inf = fminf((pre ? pre->infil : gld(I.infil, k)) * dt, d + rr);
  }
  float4 ga = make_float4(0.f, 0.f, 0.f, 0.f), gi = make_float4(0.f, 0.f, 0.f, 0.f);
  float soilIn = 0.f;
  if constexpr (EXT) {
    float cap = (I.flags & FLAG_INFIL) ? gld(I.infil, k) * dt : 0.f;
    if (I.soil) {
      ga = gld(I.soil, k);
      if (ga.z > 0.f) cap = ga.w >= ga.y ? 0.f : fminf(cap * (1.f + ga.x / fmaxf(ga.w, 1e-3f)), (ga.y - ga.w) * ga.z);
    }
    inf = fminf(cap, d + rr);
    soilIn = inf;
    if (I.gi) {
      gi = gld(I.gi, k);
      if (gi.y > 0.f) {
        const float st = fmaxf(0.f, gi.w - gi.z * dt);
        const float g2 = fminf(fminf(gi.x * dt, d + rr - inf), gi.y - st);
        gi.w = st + fmaxf(0.f, g2);
        if (g2 > 0.f) inf += g2;
      }
    }
  }
  float nh = qh, nw = qw;
  kadd(nh, nw, rr - inf);
  if (nh < 0.f) { inf = d - qw + rr; qh = 0.f; qw = 0.f; }
  else { qh = nh; qw = nw; }
  if constexpr (EXT) {
    // the soil takes its share of what actually left the cell (FP32 rounding can clip the sum)
    if (I.soil && ga.z > 0.f) {
      ga.w += fminf(soilIn, inf) / ga.z;
      gst(reinterpret_cast<fs_v4 *>(I.soil), k, fs_v4{ga.x, ga.y, ga.z, ga.w});
    }
    if (I.gi && gi.y > 0.f) gst(reinterpret_cast<fs_v4 *>(I.gi), k, fs_v4{gi.x, gi.y, gi.z, gi.w});
  }
  acc_inf += (double)inf * areaf;
  if (qh <= hdry) { qhu = 0.f; qhv = 0.f; }
  const float hm = pre ? pre->hmax : gld(I.hmax, k);
  if (qh > hm) gst(I.hmax, k, qh);
  if constexpr (SPEED) {
    const uint8_t mk = pre ? pre->mask : gld(I.mask, k);
    if (qh > hdry && !(mk & MASK_INLET)) smax = fmaxf(smax, cell_speed(qh, qhu, qhv, g));
  }
}

// Hazard products of a cell after a step of length dt ending at tnow (hazard, hzRecord,
// kernels.wgsl.js:415-447; the classes of hazard.js:30-66): the worst ARR 2019 H1-H6 and FD2320
// classes from depth and speed at the same moment, peak d*v and force per width, time deeper than
// 15 cm and its first time, peak speed and the first time deeper than VCUT. Films under 1 cm are not
// recorded; speed counts only from VCUT deep.
constexpr float kHzWet = 0.01f, kHzVcut = 0.05f, kHzThr = 0.15f;
__device__ __forceinline__ float hz_arr(float d, float v) {
  const float dv = d * v;
  if (dv <= 0.3f && d <= 0.3f && v <= 2.f) return 1.f;
  if (dv <= 0.6f && d <= 0.5f && v <= 2.f) return 2.f;
  if (dv <= 0.6f && d <= 1.2f && v <= 2.f) return 3.f;
  if (dv <= 1.f && d <= 2.f && v <= 2.f) return 4.f;
  if (dv <= 4.f && d <= 4.f && v <= 4.f) return 5.f;
  return 6.f;
}
__device__ __forceinline__ float hz_hr(float hr) { return hr < 0.75f ? 1.f : hr < 1.25f ? 2.f : hr < 2.f ? 3.f : 4.f; }

__device__ __forceinline__ void hazard_cell(const DevInst &I, int k, float d, float hu, float hv, float dt,
                                         float tnow) {
  if (!(d >= kHzWet)) return;
  const float v = d >= kHzVcut ? sqrtf(hu * hu + hv * hv) / d : 0.f;
  float4 a = gld(I.hzA, k), b = gld(I.hzB, k);
  a.x = fmaxf(a.x, d * v);
  a.y = fmaxf(a.y, d * v * v + 0.5f * I.g * d * d);
  if (d > kHzThr) { a.z += dt; if (a.w < 0.f) a.w = tnow; }
  if (d > kHzVcut) { b.x = fmaxf(b.x, v); if (b.y < 0.f) b.y = tnow; }
  const float hr = d * (v + 0.5f);
  const int packed = (int)b.z;
  const int c0 = max(packed & 7, (int)hz_arr(d, v));
  const int c1 = max((packed >> 3) & 7, (int)hz_hr(hr));
  const int c2 = max(packed >> 6, (int)hz_hr(hr + (d > 0.25f ? 1.f : 0.f)));
  b.z = (float)(c0 + 8 * c1 + 64 * c2);
  gst(reinterpret_cast<fs_v4 *>(I.hzA), k, fs_v4{a.x, a.y, a.z, a.w});
  gst(reinterpret_cast<fs_v4 *>(I.hzB), k, fs_v4{b.x, b.y, b.z, b.w});
}

// wavefront index for a device-side count of n work items: the blocks that have work are spread
// round-robin over the 8 XCDs and each XCD takes one contiguous run of items; -1 = no work
__device__ __forceinline__ int xcd_wave_index_n(int n, int nwaves_per_block) {
  const int b = blockIdx.x, x = b & 7, k = b >> 3;
  const int need = (n + nwaves_per_block - 1) / nwaves_per_block;
  const int q = need >> 3, r = need & 7;
  if (k >= q + (x < r ? 1 : 0)) return -1;
  const int lb = x * q + min(x, r) + k;
  const int w = lb * nwaves_per_block + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  return w < n ? w : -1;
}

// a face's mass and momentum fluxes and the pressure corrections of its two sides (under local time stepping
// zero for a face the record does not evaluate: its lower-level cell is not at the record's level)
struct FaceRow { float F0, F1, F2, cL, cR; };

// One strip of one RK stage (see the file comment for the row pipeline).
// STAGE 0: U* = U^n + dt L(U^n)                   (updateAt STAGE 0, kernels.wgsl.js:489-492)
// STAGE 1: U^{n+1} = (U^n + U* + dt L(U*)) / 2, then friction, rain, infiltration, max depth,
//          and the wave speed of the result for the next step's dt (kernels.wgsl.js:493-498, 544-590)
// LTS: a record of level m (solver-cpu.js:529-558; kernels.wgsl.js:313-324, 504-537): only faces
// whose lower-level cell is at m pass flux; cells at m step from their state plus their gathered
// store; cells at m + 1 wait and gather the RK-weighted flux of those faces; the bed-slope term of a
// stepping cell takes, for a face whose neighbour is finer, its reconstruction from that face's last
// evaluation (kept in bsx/bsy by the finer record's last stage).
// EXT: extended faces and edges, for instances with the step-flux rules (F26; global stepping only), flood
// barriers (crests in the face beds, weir faces, the barrier faces' fluxes for their books), stage cells,
// Green-Ampt soil or GI stores, implicit friction (F26; global stepping: under LTS every variant takes it)
// or the van Leer limiter
// IN: a strip whose every lane is an interior column and whose rows, halo rows included, are interior rows:
// no ghost cells and no domain-edge faces (the common case, compiled as a loop of its own)
template <int STAGE, bool LTS, bool HZ, bool EXT = false, bool IN = false>
__device__ __forceinline__ bool run_strip(const DevInst &I, const int4 wk, const int lane, const int m, const float dt,
                                          Part *__restrict__ part, const int cells, const float tnow,
                                          const bool acc = false) {
  constexpr bool SF = EXT;
  // the plain variants issue the pass's later loads with the row's (see below): both stages of global stepping
  // and the first of local time stepping (FS_PRE bits 0, 3 and 1; bit 2, the LTS second stage, spills registers
  // at four waves per SIMD). Global first stage, measured at 8192^2: 0.925 ms without, 0.859 ms with.
  constexpr bool PRE = !EXT && !HZ && ((!LTS && STAGE == 1 && (FS_PRE & 1)) || (LTS && STAGE == 0 && (FS_PRE & 2)) ||
                                       (LTS && STAGE == 1 && (FS_PRE & 4)) || (!LTS && STAGE == 0 && (FS_PRE & 8)));
  const bool sf = EXT && (I.flags & FLAG_STEPF);
  const bool bar = EXT && I.crest != nullptr;
  const float rain = I.rec->rainRate;
  const float g = I.g, hdry = I.hdry, hfront = I.hfront, inv_dx = I.inv_dx, dx = I.dx;
  // the van Leer limiter runs in the extended variant only: the plain kernels are minmod at compile time
  const int so = I.secondOrder, vl = EXT ? I.limiter : 0;
  const int nx = I.nx, ny = I.ny, pitch = I.pitch;
  const int X = wk.y - I.hl + lane;
  const bool own_col = lane >= I.hl && lane < I.hl + I.own && (IN || X < nx);
  const int y0 = wk.z, y1 = min(wk.z + wk.w, ny);
  const bool last = STAGE == 1 || !so;   // the stage that ends the step (or the record)
  const bool xin = IN || (wk.y - I.hl >= 0 && wk.y - I.hl + kLanes <= nx);  // every lane an interior column
  const bool xedge = !IN && (wk.y - I.hl <= 0 || wk.y - I.hl + kLanes >= nx);  // a lane on the first or last column

  // writes go to B in stage 0 and to A in stage 1; Ws is the residual of the evaluated state
  const float *__restrict__ Ws = STAGE == 0 ? I.w : I.w1;
  float *__restrict__ Ho = STAGE == 0 ? I.h1 : I.h;
  float *__restrict__ HUo = STAGE == 0 ? I.hu1 : I.hu;
  float *__restrict__ HVo = STAGE == 0 ? I.hv1 : I.hv;
  float *__restrict__ Wo = STAGE == 0 ? I.w1 : I.w;

  const float hg = 0.5f * g;
  const float wgt = so ? 0.5f : 1.f;
  const float areaf = dx * dx;

  // ---- register window (rows r-1, r-2, r-3 of the walk and what the later sections carry)
  float h1r = 0, u1r = 0, v1r = 0, z1r = 0, h2r = 0, u2r = 0, v2r = 0, z2r = 0;
  bool M1r = false, M2r = false;
  bool R1r = false, R2r = false;   // EXT: Manning n > 0 in rows r-1, r-2
  float cy1 = -1e30f, wy1 = 0.f;   // EXT: crest and weir coefficient of row r-1's south faces
  float cx1 = -1e30f, wx1 = 0.f, cx2 = -1e30f, wx2 = 0.f;   // EXT: the same of the west faces of rows r-1, r-2
  int lv1 = 0, lv2 = 0, lv3 = 0, lv4 = 0;
  float hu1 = 0, hv1 = 0, hu2 = 0, hv2 = 0, hu3 = 0, hv3 = 0, hh3 = 0;
  bool M3 = false, in1 = false, in2 = false, in3 = false;
  float yp_h = 0, yp_e = 0, yp_n = 0, yp_t = 0;
  FaceRow xf3 = {0, 0, 0, 0, 0};                          // x faces of row r-3
  FaceRow yf3 = {0, 0, 0, 0, 0}, yf4 = {0, 0, 0, 0, 0};   // y faces (r-3 | r-2) and (r-4 | r-3)
  float th3 = 1.f, th4 = 1.f;
  float st3 = 0.f;   // LTS first stage: the store of row r-3's cell (theta's load of the previous pass)
  // bed-slope terms of row r-3 (x) and rows r-2, r-3 (y); under LTS a stepping cell's term already
  // holds, for a face whose neighbour is finer, that face's last evaluation
  float sx3 = 0, sy2 = 0, sy3 = 0;

  double acc_out = 0.0, acc_clip = 0.0, acc_inf = 0.0;
  float smax = 0.f;
  bool stir = false;   // an owned cell ends with h, w, hu or hv non-zero

  // LTS second stage: which state a cell of row r is read from depends on its tile level, so the levels
  // of the next row are loaded a row ahead (one byte a lane) and the row's loads wait on nothing.
  // The load is issued first in each pass (rows and columns clamped into the grid, so it needs no branch):
  // waiting on the row's own loads then covers it.
  constexpr bool LVA = LTS && STAGE == 1;
  const int xlv = (min(max(X, 0), nx - 1)) >> 3;
  int lvAhead = 0;
  if constexpr (LVA) lvAhead = gld(I.lv, (min(max(y0 - kHalo, 0), ny - 1) >> 3) * I.tx + xlv);

  // One pass of the walk: row r loaded; y reconstruction of row r-1 and y face (r-2 | r-1); x reconstruction
  // and x faces of row r-2 (from the window: the x faces then live one pass, not three); theta of row r-2;
  // update of row r-3. The strip needs x faces and theta of rows y0-1 .. y1, y reconstructions of rows
  // y0-2 .. y1+1 and y faces from (y0-2 | y0-1) on, so the first four passes are peeled off with the
  // sections they need: PH 0 and 1 load only, PH 2 adds the y reconstruction, PH 3 the y face; PH 4 (every
  // later pass) runs all sections.
  auto pass = [&](const int r, auto phase) {
    constexpr int PH = decltype(phase)::value;
    // ---------------- row r: primitives
    const int lvRow = lvAhead;
    if constexpr (LVA) {
      lvAhead = gld(I.lv, (min(max(r + 1, 0), ny - 1) >> 3) * I.tx + xlv);
      __builtin_amdgcn_sched_barrier(0);   // (kept ahead of the row's loads)
    }
    Prim p;
    if (IN || (xin && r >= 0 && r < ny)) p = load_prim_row<STAGE, LTS, EXT>(I, X, r, m, dt, lvRow);
    else p = load_prim<STAGE, LTS, EXT>(I, X, r, m, dt);

    // ---------------- the pass's other loads, issued here in the order it consumes them: theta's store of
    // row r-2 (LTS), then the update's operands of row r-3 (its residual, RK start state and store), then its
    // sources' layers. The memory counter drains in issue order, so loads issued at their sections would
    // each wait on their own; issued together, one wait covers them and their latency overlaps the faces.
    float tsh = 0.f, tsw = 0.f, tst = 0.f;
    // LTS: row r-2's cell waits beside a cell at m. Only then does its theta scale a flux: every face of a
    // waiting cell without a neighbour at m is inactive this record, so its store need not be read.
    // (Lane moves run on every lane, outside any condition on a lane's own data: inside such a branch a move
    // reads the inactive lanes as zero, and the compiler may stand in the branch condition's value for the
    // moved operand. So no lane move sits in the right operand of && or ||.)
    bool wb2 = false;
    if constexpr (LTS) {
      const int lvL = from_left_i(lv2), lvR = from_right_i(lv2);
      wb2 = (lv2 == m + 1) & ((lvL == m) | (lvR == m) | (lv1 == m) | (lv3 == m));
    }
    float pw = 0.f, ph0 = 0.f, pw0 = 0.f, phu0 = 0.f, phv0 = 0.f, psh = 0.f, psw = 0.f, pshu = 0.f, pshv = 0.f;
    SrcOps pso = {0.f, 0.f, 0.f, 1.f, (uint8_t)0};
    if constexpr (PRE && PH == 4) {
      if constexpr (LTS) {
        if (in2 && M2r && (wb2 || (STAGE == 0 && lv2 == m))) {
          const int k2 = (r - 2) * pitch + X;
          if (STAGE == 0) { const float2 t = gld_lo(I.st4, k2); tsh = t.x; tsw = t.y; }
          else tst = gld(reinterpret_cast<const float *>(I.g4), 4 * k2 + 3);
        }
      }
      const int ry = r - kHalo;
      bool take = own_col && ry >= y0 && ry < y1 && in3 && M3;
      if constexpr (LTS) take = take && lv3 == m;
      if (take) {
        const int k = ry * pitch + X;
        pw = gld(Ws, k);
        if (STAGE == 1) { ph0 = gld(I.h, k); pw0 = gld(I.w, k); phu0 = gld(I.hu, k); phv0 = gld(I.hv, k); }
        if constexpr (LTS) { const float4 t = gld(I.st4, k); psh = t.x; psw = t.y; pshu = t.z; pshv = t.w; }
        if (STAGE == 1) pso = load_src<!LTS>(I, k);
      }
    }
    // barrier crests of this row: (west face crest, coefficient, south face crest, coefficient)
    float4 cr = make_float4(-1e30f, 0.f, -1e30f, 0.f);
    if constexpr (EXT) {
      if (bar && (IN || (r >= 0 && r < ny && X >= 0 && X < nx))) cr = gld(I.crest, r * pitch + X);
    }

    // ---------------- y reconstruction of row r-1 (neighbours: rows r-2 and r in registers)
    FaceRow yf2 = {0.f, 0.f, 0.f, 0.f, 0.f};
    float syr = 0.f;   // y bed slope of row r-1 (moves on at the end of the pass: the update uses row r-3's)
    if constexpr (PH >= 2) {
      const Slopes sys = muscl<SF>(h1r, v1r, u1r, h2r, v2r, u2r, M2r, z1r - z2r, p.h, p.v, p.u, p.M, p.z - z1r, so,
                                   hfront, vl, sf && R1r);
      const float yh_p = h1r + 0.5f * sys.sh, ye_p = h1r + 0.5f * sys.se;
      const float yh_m = h1r - 0.5f * sys.sh, ye_m = h1r - 0.5f * sys.se;
      const float yn_p = v1r + 0.5f * sys.sn, yt_p = u1r + 0.5f * sys.st;
      const float yn_m = v1r - 0.5f * sys.sn, yt_m = u1r - 0.5f * sys.st;
      const float sy1 = hg * (yh_m + yh_p) * ((ye_p - yh_p) - (ye_m - yh_m));
      const float4 ry1 = {yh_m, ye_m - yh_m, yh_p, ye_p - yh_p};
      // y face (r-2 | r-1): L = row r-2 toward north, R = row r-1 toward south
      bool onS = true;   // LTS: the record evaluates the y face (r-2 | r-1)
      if constexpr (PH >= 3) {
        bool weirY = false;
        if constexpr (EXT) weirY = bar && M2r && M1r && cy1 > fmaxf(z2r, z1r);
        Face f = face<EXT>(yp_h, yp_e, yp_n, yp_t, M2r, yh_m, ye_m, yn_m, yt_m, M1r, z1r - z2r, g, hdry,
                           sf && R2r && !weirY, R1r, cy1 - z2r);
        if constexpr (EXT) {
          if (weirY) weir_flux(f, h2r + z2r, h2r, u2r, h1r + z1r, h1r, u1r, cy1, wy1, g, hdry, dx, dt);
        }
        yf2 = FaceRow{f.F0, f.F1, f.F2, f.cL, f.cR};
        if constexpr (LTS) {
          onS = min(lv2, lv1) == m;
          if (!onS) yf2 = FaceRow{0.f, 0.f, 0.f, 0.f, 0.f};
        }
      }
      syr = sy1;
      if constexpr (LTS && PH == 4) {   // (row r-1 is an owned row only from the first full pass on)
        const int ry = r - 1;
        const bool onN = min(lv1, p.lv) == m;
        const bool row_in = own_col && in1 && M1r && ry >= y0 && ry < y1;
        if (last && row_in && lv1 == m + 1) {
          float2 *b = reinterpret_cast<float2 *>(I.bsy);
          const int kb = 2 * (ry * pitch + X);
          if (onS) gst(b, kb, make_float2(ry1.x, ry1.y));
          if (onN) gst(b, kb + 1, make_float2(ry1.z, ry1.w));
        }
        if (row_in && lv1 == m && (!onS || !onN)) {
          const float4 b = gld(I.bsy, ry * pitch + X);
          float4 ys = ry1;
          if (!onS) { ys.x = b.x; ys.y = b.y; }
          if (!onN) { ys.z = b.z; ys.w = b.w; }
          syr = hg * (ys.x + ys.z) * (ys.w - ys.y);
        }
      }
      yp_h = yh_p; yp_e = ye_p; yp_n = yn_p; yp_t = yt_p;
    }

    if constexpr (PH == 4) {
      // ---------------- x reconstruction of row r-2 (neighbours by lane shifts)
      const int rx = r - 2;
      const float hl = from_left(h2r), ul = from_left(u2r), vlf = from_left(v2r), zl = from_left(z2r);
      const float hr = from_right(h2r), ur = from_right(u2r), vr = from_right(v2r), zr = from_right(z2r);
      const bool Ml = from_left_i(M2r) != 0, Mr = from_right_i(M2r) != 0;
      bool Rr = false;
      if constexpr (SF) Rr = from_right_i(R2r) != 0;
      const Slopes sxs = muscl<SF>(h2r, u2r, v2r, hl, ul, vlf, Ml, z2r - zl, hr, ur, vr, Mr, zr - z2r, so, hfront, vl,
                                   sf && R2r);
      const float xh_p = h2r + 0.5f * sxs.sh, xe_p = h2r + 0.5f * sxs.se;
      const float xh_m = h2r - 0.5f * sxs.sh, xe_m = h2r - 0.5f * sxs.se;
      const float xu_p = u2r + 0.5f * sxs.sn, xv_p = v2r + 0.5f * sxs.st;
      const float xu_m = u2r - 0.5f * sxs.sn, xv_m = v2r - 0.5f * sxs.st;
      const float sx0 = hg * (xh_m + xh_p) * ((xe_p - xh_p) - (xe_m - xh_m));
      const float4 rx0 = {xh_m, xe_m - xh_m, xh_p, xe_p - xh_p};
      // x face (l | l+1): right side is lane l+1's minus side
      FaceRow xf2;
      {
        const float Rh = from_right(xh_m), Re = from_right(xe_m), Ru = from_right(xu_m), Rv = from_right(xv_m);
        float zcx = -1e30f, cwx = 0.f;   // crest of this face (held by lane l + 1 as its west face)
        bool weirX = false;
        if constexpr (EXT) {
          zcx = from_right(cx2);
          cwx = from_right(wx2);
          weirX = bar && M2r && Mr && zcx > fmaxf(z2r, zr);
        }
        Face f = face<EXT>(xh_p, xe_p, xu_p, xv_p, M2r, Rh, Re, Ru, Rv, Mr, zr - z2r, g, hdry, sf && R2r && !weirX, Rr,
                           zcx - z2r);
        if constexpr (EXT) {
          if (weirX) weir_flux(f, h2r + z2r, h2r, v2r, hr + zr, hr, vr, zcx, cwx, g, hdry, dx, dt);
        }
        xf2 = FaceRow{f.F0, f.F1, f.F2, f.cL, f.cR};
      }
      bool onE = true;   // LTS: the record evaluates this x face (l | l+1)
      if constexpr (LTS) {
        onE = min(lv2, from_right_i(lv2)) == m;
        if (!onE) xf2 = FaceRow{0.f, 0.f, 0.f, 0.f, 0.f};
      }
      float sx2 = sx0;
      if constexpr (LTS) {
        // a waiting cell keeps its reconstruction toward the faces this record evaluates (last stage)
        const bool onW = from_left_i(onE ? 1 : 0) != 0;
        const bool row_in = own_col && in2 && M2r && rx >= y0 && rx < y1;
        if (last && row_in && lv2 == m + 1) {
          float2 *b = reinterpret_cast<float2 *>(I.bsx);
          const int kb = 2 * (rx * pitch + X);
          if (onW) gst(b, kb, make_float2(rx0.x, rx0.y));
          if (onE) gst(b, kb + 1, make_float2(rx0.z, rx0.w));
        }
        // a stepping cell's bed slope from its face reconstructions (solver-cpu.js:363-368): a face with
        // a finer neighbour contributes the values of its last evaluation
        if (row_in && lv2 == m && (!onW || !onE)) {
          const float4 b = gld(I.bsx, rx * pitch + X);
          float4 xs = rx0;
          if (!onW) { xs.x = b.x; xs.y = b.y; }
          if (!onE) { xs.z = b.z; xs.w = b.w; }
          sx2 = hg * (xs.x + xs.z) * (xs.w - xs.y);
        }
      }
      // ---------------- theta of row r-2 (solver-cpu.js:307-318; kernels.wgsl.js:349-363)
      float th2 = 1.f;
      float st2 = 0.f;   // LTS first stage: the store of row r-2's cell (its gather uses it in the next pass)
      {
        const float fe = xf2.F0, fw = from_left(xf2.F0), fs = yf3.F0, fn = yf2.F0;
        bool take = in2 && M2r;
        float hv_ = h2r;
        const float out = ((fe > 0.f ? fe : 0.f) - (fw < 0.f ? fw : 0.f) + (fn > 0.f ? fn : 0.f) - (fs < 0.f ? fs : 0.f)) * dx * dt;
        if constexpr (LTS) {
          take = take && (lv2 == m || lv2 == m + 1);
          // a waiting cell, or a stepping one in its first stage, holds its water plus its store (the second
          // stage reads a waiting cell's only if it passes water out: else theta stays 1 whatever it holds; the
          // first stage's read also makes the snapshot the cell's gather leaves for the second stage)
          if (take && (lv2 == m ? STAGE == 0 : wb2 && (STAGE == 0 || out > 0.f))) {
            const int k2 = (r - 2) * pitch + X;
            float st;
            if (STAGE == 0) {
              if constexpr (PRE) st = tsh - tsw;
              else { const float2 t = gld_lo(I.st4, k2); st = t.x - t.y; }
              st2 = st;
            } else {
              // a waiting cell in the second stage: its store with the first stage's gather, as the first
              // stage left it (the cell's own wavefront folds the record's gathers into the store meanwhile)
              st = PRE ? tst : gld(reinterpret_cast<const float *>(I.g4), 4 * k2 + 3);
            }
            hv_ = fmaxf(0.f, h2r + st);
          }
        }
        const float vol = hv_ * areaf;
        if (take && out > vol) th2 = vol / out;
      }

      // ---------------- update of row r-3
      {
        const int ry = r - kHalo;
        // west face (l-1 | l) from lane l-1; east face (l | l+1) own
        const float Fw0 = from_left(xf3.F0), Fw1 = from_left(xf3.F1), Fw2 = from_left(xf3.F2), cRw = from_left(xf3.cR);
        const float thW = from_left(th3), thE = from_right(th3);
        float Fe0 = xf3.F0, Fe1 = xf3.F1, Fe2 = xf3.F2;
        const float cLe = xf3.cL;
        float Fs0 = yf4.F0, Fs1 = yf4.F1, Fs2 = yf4.F2;
        const float cRs = yf4.cR;
        float Fn0 = yf3.F0, Fn1 = yf3.F1, Fn2 = yf3.F2;
        const float cLn = yf3.cL;
        // each face carries its donor's theta (solver-cpu.js:319-326; kernels.wgsl.js:364-365)
        const float tw = Fw0 > 0.f ? thW : th3;
        const float te = Fe0 > 0.f ? th3 : thE;
        const float ts = Fs0 > 0.f ? th4 : th3;
        const float tn = Fn0 > 0.f ? th3 : th2;
        const float w0 = Fw0 * tw, w1 = Fw1 * tw, w2 = Fw2 * tw;
        Fe0 *= te; Fe1 *= te; Fe2 *= te;
        Fs0 *= ts; Fs1 *= ts; Fs2 *= ts;
        Fn0 *= tn; Fn1 *= tn; Fn2 *= tn;
        if constexpr (EXT) {
          // a barrier on this cell's west or south face: its theta-scaled mass flux for the books (one lane
          // owns the face, so a plain store)
          if (bar && own_col && ry >= y0 && ry < y1 && in3) {
            const int2 bi = gld(I.bidx, ry * pitch + X);
            if (bi.x >= 0) I.barFlux[bi.x] = w0;
            if (bi.y >= 0) I.barFlux[bi.y] = Fs0;
          }
        }
        bool upd = own_col && ry >= y0 && ry < y1 && in3 && M3;
        bool own = true, coarse = false;
        // LTS: a waiting cell gathers when one of its faces is evaluated, i.e. a face neighbour is at m
        bool gathers = false;
        if constexpr (LTS) {
          const int lvL = from_left_i(lv3), lvR = from_right_i(lv3);   // (every lane: see wb2)
          gathers = (lvL == m) | (lvR == m) | (lv2 == m) | (lv4 == m);
        }
        if constexpr (LTS) {
          own = lv3 == m;
          coarse = lv3 == m + 1;
          upd = upd && (own || coarse);
        }
        if (upd) {
          const int k = ry * pitch + X;
          float bx = sx3, by = sy3;
          if constexpr (LTS) {
            if (!own) { bx = 0.f; by = 0.f; }   // a waiting cell gathers fluxes only
          }
          const float dh = -inv_dx * (Fe0 - w0 + Fn0 - Fs0);
          const float dhu = -inv_dx * ((Fe1 + cLe) - (w1 + cRw) + bx + Fn2 - Fs2);
          const float dhv = -inv_dx * ((Fn1 + cLn) - (Fs1 + cRs) + by + Fe2 - w2);
          if (LTS && coarse) {
            // waiting cell: gather the RK-weighted change (ltsUpdate mode 2, kernels.wgsl.js:515-519).
            // First stage: the change into g0 (neighbouring strips read this cell's store for theta during the
            // stage, so it stays), and the store the second stage's theta reads, the first stage's change
            // included, into tst. Second stage: the cell's own wavefront adds both stages' changes to its store
            // (compensated, as kadd on U0), which no other wavefront reads during that stage. First order: the
            // change and a flag, added after the record (sw_lts_fold).
            if (gathers) {
              const float wd = so ? 0.5f * dt : dt;
              const float g0 = mul_nc(wd, dh), g0u = mul_nc(wd, dhu), g0v = mul_nc(wd, dhv);
              if (STAGE == 0) {
                gst(I.g4, k, make_float4(g0, g0u, g0v, so ? add_nc(st3, g0) : 0.f));
                if (!so) gst(I.gflag, k, (uint8_t)1);
              } else {
                const float4 t = gld(I.st4, k);
                float sh = t.x, sw = t.y, shu = t.z, shv = t.w;
                const float4 ga = gld(I.g4, k);
                const float a0 = ga.x, a0u = ga.y, a0v = ga.z;
                kadd(sh, sw, a0);
                shu = add_nc(shu, a0u);
                shv = add_nc(shv, a0v);
                kadd(sh, sw, g0);
                shu = add_nc(shu, g0u);
                shv = add_nc(shv, g0v);
                gst(I.st4, k, make_float4(sh, sw, shu, shv));
              }
            }
          } else {
            float qh = hh3, qhu = hu3, qhv = hv3, qw = PRE ? pw : gld(Ws, k);
            // RK start state: U^n (the window's values in the first stage), plus (LTS) what finer
            // neighbours passed while the cell waited
            float h0 = 0.f, w0n = 0.f, hu0 = 0.f, hv0 = 0.f;
            bool spent = false;   // LTS: the cell's store held gathered flux (cleared once used)
            if (STAGE == 1 || LTS) {
              if (STAGE == 1) {
                if constexpr (PRE) { h0 = ph0; w0n = pw0; hu0 = phu0; hv0 = phv0; }
                else {
                  h0 = gld(I.h, k); w0n = gld(I.w, k); hu0 = gld(I.hu, k); hv0 = gld(I.hv, k);
                  if constexpr (!LTS && !EXT) pso = load_src(I, k);   // (as below, under LTS)
                }
              } else {
                h0 = hh3; w0n = qw; hu0 = hu3; hv0 = hv3;
              }
              if constexpr (LTS) {
                float sh = psh, sw = psw, shu = pshu, shv = pshv;
                if constexpr (!PRE) { const float4 t = gld(I.st4, k); sh = t.x; sw = t.y; shu = t.z; shv = t.w; }
                // the sources' layers, issued after the update's operands so that one wait covers them
                // (sources_cell would otherwise wait on each in turn)
                if constexpr (STAGE == 1 && !PRE && !EXT) pso = load_src<!LTS>(I, k);
                hu0 += shu;
                hv0 += shv;
                kadd(h0, w0n, sh - sw);
                spent = sh != 0.f || sw != 0.f || shu != 0.f || shv != 0.f;
              }
            }
            if (STAGE == 0) {
              if constexpr (LTS) { qh = h0; qw = w0n; qhu = hu0; qhv = hv0; }
              qhu = qhu + dt * dhu;
              qhv = qhv + dt * dhv;
              kadd(qh, qw, dt * dh);
            } else {
              qhu = 0.5f * (hu0 + qhu + dt * dhu);
              qhv = 0.5f * (hv0 + qhv + dt * dhv);
              // Heun average of the true depths (h - w) as one compensated increment (kernels.wgsl.js:
              // 493-498), with the stored depths differenced first and the residuals kept apart:
              // (h0 - w0) - (h* - w*) rounded at the depth's ulp drops the residuals every step, a drift
              // of 1e-6 of the volume over 10^4 steps of a point inflow
              kadd(qh, qw, 0.5f * (((h0 - qh) + (qw - w0n)) + dt * dh));
            }
            if (qh < 0.f) { acc_clip += -(double)(qh - qw) * areaf; qh = 0.f; qw = 0.f; }
            if (qh <= hdry) { qhu = 0.f; qhv = 0.f; }
            // boundary outflow through the domain edge faces (edgeOutAt, kernels.wgsl.js:1020-1030)
            // (only strips holding a domain edge column and the edge rows take this wave-uniform branch)
            if (!IN && (xedge || ry == 0 || ry == ny - 1)) {
              double o = 0.0;
              if (X == 0) o -= (double)w0;
              if (X == nx - 1) o += (double)Fe0;
              if (ry == 0) o -= (double)Fs0;
              if (ry == ny - 1) o += (double)Fn0;
              acc_out += o * ((double)dx * (double)dt * (double)wgt);
            }
            if (STAGE == 1) {
              sources_cell<EXT, EXT || LTS, !LTS>(I, k, dt, rain, areaf, qh, qw, qhu, qhv, acc_inf, smax, EXT ? nullptr : &pso);
              if constexpr (HZ) {
                if (I.hzA) hazard_cell(I, k, qh, qhu, qhv, dt, tnow);
              }
              if constexpr (LTS) {
                if (spent) {
                  gst(I.st4, k, make_float4(0.f, 0.f, 0.f, 0.f));
                }
              }
            }
            gst(Ho, k, qh); gst(HUo, k, qhu); gst(HVo, k, qhv); gst(Wo, k, qw);
            stir = stir || qh != 0.f || qw != 0.f || qhu != 0.f || qhv != 0.f;
          }
        }
      }

      // ---------------- x faces, theta and bed slope of row r-2 move on
      yf4 = yf3;
      th4 = th3; th3 = th2;
      if constexpr (LTS && STAGE == 0) st3 = st2;
      xf3 = xf2;
      sx3 = sx2;
    }
    if constexpr (PH >= 3) yf3 = yf2;   // (yf4 moved on above; before PH 4 the faces are not needed)
    if constexpr (PH >= 2) { sy3 = sy2; sy2 = syr; }
    // ---------------- rotate the rows
    if constexpr (LTS) { lv4 = lv3; lv3 = lv2; lv2 = lv1; lv1 = p.lv; }
    hh3 = h2r; hu3 = hu2; hv3 = hv2; M3 = M2r; in3 = in2;
    h2r = h1r; u2r = u1r; v2r = v1r; z2r = z1r; M2r = M1r; hu2 = hu1; hv2 = hv1; in2 = in1;
    h1r = p.h; u1r = p.u; v1r = p.v; z1r = p.z; M1r = p.M; hu1 = p.hu; hv1 = p.hv; in1 = p.in;
    if constexpr (EXT) {
      R2r = R1r; R1r = p.R; cy1 = cr.z; wy1 = cr.w;
      cx2 = cx1; wx2 = wx1; cx1 = cr.x; wx1 = cr.y;
    }
  };
  pass(y0 - 3, std::integral_constant<int, 0>());
  pass(y0 - 2, std::integral_constant<int, 1>());
  pass(y0 - 1, std::integral_constant<int, 2>());
  pass(y0, std::integral_constant<int, 3>());
  // (unrolled by two, the global kernels issue 6 % fewer VALU instructions but step only 0.7 % faster at
  // 8192^2, and the compiler fuses multiply-adds differently in the two copies of a pass: the state then
  // differs in its last bits from the single-pass loop's)
  for (int r = y0 + 1; r < y1 + kHalo; ++r) pass(r, std::integral_constant<int, 4>());

  const double o = wave_sum(acc_out), c = wave_sum(acc_clip), f = wave_sum(acc_inf);
  const float s = wave_max(smax);
  if (lane == 0) {
    if (acc) {   // a further part of the same wavefront (LTS): its books join the partial
      Part q = *part;
      q.out += o; q.clip += c; q.infil += f; q.smax = fmaxf(q.smax, s); q.cells += cells;
      *part = q;
    } else {
      *part = Part{o, c, f, s, cells};
    }
  }
  return __any(stir);
}

// a work item whose strip, halo rows and columns included, lies inside the grid
__device__ __forceinline__ bool strip_interior(const DevInst &I, const int4 wk) {
  return wk.y - I.hl >= 0 && wk.y - I.hl + kLanes <= I.nx && wk.z >= kHalo && min(wk.z + wk.w, I.ny) + kHalo <= I.ny;
}

template <int STAGE, bool LTS, bool HZ, bool EXT>
__device__ __forceinline__ bool run_strip_any(const DevInst &I, const int4 wk, const int lane, const int m, const float dt,
                                              Part *__restrict__ part, const int cells, const float tnow,
                                              const bool acc = false) {
  if (strip_interior(I, wk)) return run_strip<STAGE, LTS, HZ, EXT, true>(I, wk, lane, m, dt, part, cells, tnow, acc);
  return run_strip<STAGE, LTS, HZ, EXT, false>(I, wk, lane, m, dt, part, cells, tnow, acc);
}

// ------------------------------------------------------------ dry strips (global stepping)
// advance strip u this step: rain falls on it, or a strip of its 3x3 block is not quiet
// (the stencil of a whole step reaches 6 cells, less than a strip's width and height)
__device__ __forceinline__ bool strip_active(const DevInst &I, int u) {
  // srain bit 0: an open cell with rain; bit 1: a stage cell (always stepping, like the reference's
  // stage-edge tiles, solver-gpu.js:134-143)
  const uint8_t sr = I.srain[u];
  if ((sr & 2) || (I.rec->rainRate > 0.f && (sr & 1))) return true;
  const int sx = I.sxn, sy = (I.wend - I.wbeg) / I.sxn;
  const int ux = u % sx, uy = u / sx;
  bool any = false;
  for (int j = max(uy - 1, 0); j <= min(uy + 1, sy - 1); ++j)
    for (int i = max(ux - 1, 0); i <= min(ux + 1, sx - 1); ++i) any = any || I.wet[j * sx + i] != 0;
  return any;
}

// first stage: decide whether strip u steps; a strip that stops stepping clears its U* once, so the
// neighbours' second stage reads the zero state a computed first stage would have left there
__device__ __forceinline__ bool strip_begin(const DevInst &I, const int4 wk, int u, int lane) {
  const bool active = strip_active(I, u);
  const bool was = I.act[u] != 0;
  if (active == was) return active;
  if (lane == 0) I.act[u] = active ? 1 : 0;
  if (!active) {
    const int X = wk.y - I.hl + lane;
    if (lane >= I.hl && lane < I.hl + I.own && X < I.nx)
      for (int y = wk.z; y < min(wk.z + wk.w, I.ny); ++y) {
        const size_t k = (size_t)y * I.pitch + X;
        I.h1[k] = 0.f; I.hu1[k] = 0.f; I.hv1[k] = 0.f; I.w1[k] = 0.f;
      }
  }
  return active;
}

// ------------------------------------------------------------ LTS record work
// A level-m record owns, in each work unit (56 columns x up to 64 rows, whole 8x8 tiles), the rows that
// hold a cell it advances or a cell that gathers: every row of a tile row with an open tile at m or an
// open tile at m + 1 whose west or east neighbour is at m (its cells next to that face gather on every
// row), and the first or last row of an open tile at m + 1 whose south or north neighbour is at m (only
// that row gathers). Tiles without open cells have nothing to advance or gather; a wall tile at m still
// makes the open tiles beside it gather (its faces pass the walls' pressure). Gaps of up to 2 kHalo rows
// between owned rows are owned too (walking them costs no more than the halo rows of a second part).
// Each run of owned rows, cut into parts of at most FS_LTS_PART rows, is one wavefront's walk of its rows
// plus kHalo halo rows each side (sw_lts_units, sw_lts_lists). A record that has few owned rows splits its
// runs into parts of 1, 2, 4, 8, 16 or 32 rows, the finest that keeps the record within kLtsFineWaves
// wavefronts, so its parts fill the chip with short walks (its latency is the longest walk). A unit holds
// at most kLtsUnitRuns parts for any part size of 8 rows or more (runs are at least 7 rows apart).
// wavefronts a record launches per work unit (the partials it has); measured at 8192^2, LTS cycle 14.66 / 13.72 /
// 13.43 / 13.50 ms with 1 / 2 / 3 / 4
constexpr int kLtsWaveCap = 3;
constexpr int kLtsFineWaves = 4096;   // parts a finely split record may have
constexpr int kLtsUnitRuns = 8;       // parts a unit can hold (64 rows, runs at least 7 apart, parts >= 8 rows)
constexpr int kLtsCand = 7;           // part sizes 1, 2, 4, 8, 16, 32 rows, then whole runs
// the most rows a part of a record that is not split finely takes (64: whole runs; measured at 8192^2, LTS cycle
// 16.68 ms with 64, 16.60 with 32, 19.11 with 16)
#ifndef FS_LTS_PART
#define FS_LTS_PART 32
#endif
constexpr int kLtsBigCand = FS_LTS_PART >= 64 ? kLtsCand - 1 : FS_LTS_PART >= 32 ? 5 : FS_LTS_PART >= 16 ? 4 : 3;

// fill the gaps of at most 2 kHalo clear bits between set bits of a unit's row mask
__device__ __forceinline__ uint64_t lts_fill_gaps(uint64_t msk) {
  uint64_t out = msk;
  int pos = 0;
  while (pos < 64) {
    const uint64_t rest = msk >> pos;
    if (!rest) break;
    const int s = pos + __builtin_ctzll(rest);                              // a run starts at s
    const uint64_t from = msk >> s;
    const int len = ~from ? __builtin_ctzll(~from) : 64 - s;                // and ends at e
    const int e = s + len;
    if (e >= 64) break;
    const uint64_t after = msk >> e;
    if (!after) break;
    const int g = __builtin_ctzll(after);                                   // clear rows before the next run
    if (g <= 2 * kHalo) out |= ((1ull << g) - 1ull) << e;
    pos = e + g;
  }
  return out;
}

// the runs of a row mask: calls f(first row, rows) for each, in row order
template <typename F>
__device__ __forceinline__ void lts_runs(uint64_t msk, F &&f) {
  int pos = 0;
  while (pos < 64) {
    const uint64_t rest = msk >> pos;
    if (!rest) break;
    const int s = pos + __builtin_ctzll(rest);
    const uint64_t from = msk >> s;
    const int len = ~from ? __builtin_ctzll(~from) : 64 - s;
    f(s, len);
    pos = s + len;
  }
}

// part size of candidate c (rows; candidate kLtsCand - 1: whole runs)
__device__ __forceinline__ int lts_cand_rows(int c) { return c < kLtsCand - 1 ? 1 << c : 64; }

// a waiting tile (m + 1, open cells) beside a tile at m: the fold's relevance test
__device__ __forceinline__ bool tile_waits(const DevInst &I, int a, int b, int m) {
  if (a < 0 || a >= I.tx || b < 0 || b >= I.ty) return false;
  const int t = b * I.tx + a;
  const uint8_t *lv = I.lv;
  if (lv[t] != m + 1 || I.tileOpen[t] == 0) return false;
  return (a > 0 && lv[t - 1] == m) || (a + 1 < I.tx && lv[t + 1] == m) || (b > 0 && lv[t - I.tx] == m) ||
         (b + 1 < I.ty && lv[t + I.tx] == m);
}

// wavefronts of a level-m record: one per part, at most kLtsWaveCap per unit (the launch grid); a wavefront
// takes parts w, w + G, w + 2 G, ... and sums their books into one partial
__device__ __forceinline__ int lts_waves(const DevInst &I, int m) {
  return min(I.rec->lcount[m], kLtsWaveCap * (I.wend - I.wbeg));
}

// end time of the record at base step kk of the current cycle (the last one ends the cycle exactly;
// solver-cpu.js:525, kernels.wgsl.js:754)
__device__ __forceinline__ float lts_record_end_time(const DevInst &I, int kk) {
  const StepRec *rec = I.rec;
  return (float)(kk + 1 == (1 << I.L) ? rec->cycT0 + rec->cycD : rec->cycT0 + (double)(kk + 1) * rec->base);
}

template <bool LTS>
__device__ __forceinline__ void exchange_inlet(const DevInst &I, const int j, const int m);

// slices of a record's partials summed apart at its end (at most)
constexpr int kLtsEndSlices = 64;

// The end of a level-m record: its books (the stages' per-wavefront partials summed in a fixed order, so the
// result does not depend on which block ends last) and the exchange of the inlets whose cell is at the level
// (exchangeL, kernels.wgsl.js:596). The partials are cut into S slices, each summed in index order into the
// slice's books; the last slice's summer (a ticket) sums the slices in slice order.

// sum slice s (partials k0 .. k1-1) into rend, then take the ticket: true for the last slice (one wavefront)
__device__ __forceinline__ bool lts_end_slice(const DevInst &I, const Part *__restrict__ p1,
                                              const Part *__restrict__ p2, int s, int k0, int k1, int S, int lane) {
  const bool two = I.secondOrder;
  double o = 0.0, c = 0.0, f = 0.0;
  for (int k = k0 + lane; k < k1; k += 64) {
    const Part a = p1[k];
    o += a.out;
    c += a.clip;
    if (two) {
      const Part q = p2[k];
      o += q.out;
      c += q.clip;
      f += q.infil;
    } else f += p2[k].infil;
  }
  o = wave_sum(o); c = wave_sum(c); f = wave_sum(f);
  unsigned t = 0;
  if (lane == 0) {
    I.rend[3 * s] = o; I.rend[3 * s + 1] = c; I.rend[3 * s + 2] = f;
    __threadfence();
    t = atomicAdd(I.ticket, 1u);
  }
  t = __shfl(t, 0, 64);
  return t == (unsigned)(S - 1);
}

// the last slice's wavefront: the record's books from the S slices, the ticket reset for the next record, the
// inlet exchange
__device__ __forceinline__ void lts_books(const DevInst &I, int m, int S, int lane) {
  __threadfence();
  StepRec *rec = I.rec;
  const float dt = rec->ldt[m];
  double O = 0.0, C = 0.0, F = 0.0;
  for (int q = lane; q < S; q += 64) {   // (fixed order: each lane's slices in turn, then a fixed tree)
    O += __hip_atomic_load(I.rend + 3 * q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    C += __hip_atomic_load(I.rend + 3 * q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    F += __hip_atomic_load(I.rend + 3 * q + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  O = wave_sum(O); C = wave_sum(C); F = wave_sum(F);
  if (lane == 0) {
    rec->outflow += O;
    rec->clipped += C;
    rec->infil += F;
    if (rec->rainCells >= 0) rec->rain += (double)(rec->rainRate * dt) * (double)rec->lcells[m] * (double)I.area;
    else rec->rain += (double)rec->rainRate * (double)dt * rec->lrf[m] * (double)I.area;
    rec->cellUpdates += (double)rec->lcells[m];
    rec->records += 1;
    rec->lastDt = dt;
  }
  if (lane == 0) __hip_atomic_store(I.ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);   // for the next record
  for (int j = lane; j < I.nin; j += 64) exchange_inlet<true>(I, j, m);
}

// LTS record of level m (single-instance groups): the parts of the level's list
template <int STAGE, bool HZ, bool EXT = false>
__device__ __forceinline__ void lts_stage_body(const DevInst *__restrict__ insts, Part *__restrict__ part, int m) {
  const int lane = threadIdx.x & 63;
  const DevInst &I = insts[0];
  if (STAGE == 1 && !I.secondOrder) return;
  const int kk = m >> 8;   // the record's base step in its cycle (second stage: for the hazard clock)
  m &= 0xff;
  const float dt = I.rec->ldt[m];
  const int np = I.rec->lcount[m], G = lts_waves(I, m);
  const int w = xcd_wave_index_n(G, kLtsWaves);
  if (w < 0) return;
  bool acc = false;
  if (dt > 0.f) {
    const float tend = lts_record_end_time(I, kk);
    for (int i = w; i < np; i += G) {
      run_strip_any<STAGE, true, HZ, EXT>(I, I.plist[m * I.capParts + i], lane, m, dt, part + w, 0, tend, acc);
      acc = true;
    }
  }
  if (!acc && lane == 0) part[w] = Part{0.0, 0.0, 0.0, 0.f, 0};
}

// One RK stage over a work list. Global stepping: one wavefront per strip of `work`; a strip the
// dry-strip rule leaves out writes an empty partial. Each kernel variant advances only the instances
// of its kind (hazard products: second stage only; extended faces: step flux or barriers) and leaves the
// others' strips alone, so a group of mixed kinds runs one launch per kind present and every instance gets
// the same machine code, hence the same bits, alone and in any mix.
__device__ __forceinline__ bool ext_faces(const DevInst &I) {
  return (I.flags & (FLAG_IMPF | FLAG_STEPF | FLAG_STAGE)) || I.crest || I.soil || I.gi || I.limiter;
}

template <int STAGE, bool LTS, bool HZ = false, bool EXT = false>
__global__ __launch_bounds__(256) void sw_stage(const DevInst *__restrict__ insts, const int4 *__restrict__ work,
                                                int nwork, Part *__restrict__ part, int m) {
  static_assert(!LTS, "local time stepping records use the specialisations below");
  const int wid = xcd_wave_index(kGlobalWaves);
  const int lane = threadIdx.x & 63;
  if (wid >= nwork) return;
  const int4 wk = work[wid];
  const DevInst &I = insts[wk.x];
  if (ext_faces(I) != EXT) return;
  if (STAGE == 1 && (I.hzA != nullptr) != HZ) return;
  if (STAGE == 1 && !I.secondOrder) return;  // first order: sw_euler_finish writes this strip's partial
  const float dt = I.rec->dt;
  const int u = wid - I.wbeg;
  bool active = dt > 0.f;
  if (active && I.wet) active = STAGE == 0 ? strip_begin(I, wk, u, lane) : I.act[u] != 0;
  if (!active) {
    if (lane == 0) part[wid] = Part{0.0, 0.0, 0.0, 0.f, 0};
    return;
  }
  const bool stir = run_strip_any<STAGE, false, HZ, EXT>(I, wk, lane, 0, dt, part + wid, I.sopen[u], (float)I.rec->tAfter);
  if (STAGE == 1 && I.wet && lane == 0) I.wet[u] = stir ? 1 : 0;
}

// waves per SIMD the LTS stage kernels are built for (the second argument of __launch_bounds__ caps
// the registers so that many fit): both stages at 4 (128 VGPRs; the second stage spills 16 bytes,
// reloaded about once per row, and runs 1.15x faster than at 3 waves with 135 VGPRs)
#ifndef FS_LTS_S1_WAVES
#define FS_LTS_S1_WAVES 4
#endif
#ifndef FS_LTS_S2_WAVES
#define FS_LTS_S2_WAVES 4
#endif
template <>
__global__ __launch_bounds__(256, FS_LTS_S1_WAVES) void sw_stage<0, true, false>(
    const DevInst *__restrict__ insts, const int4 *__restrict__ work, int nwork, Part *__restrict__ part, int m) {
  (void)nwork; (void)work;
  lts_stage_body<0, false>(insts, part, m);
}
template <>
__global__ __launch_bounds__(256, FS_LTS_S2_WAVES) void sw_stage<1, true, false>(
    const DevInst *__restrict__ insts, const int4 *__restrict__ work, int nwork, Part *__restrict__ part, int m) {
  (void)nwork; (void)work;
  lts_stage_body<1, false>(insts, part, m);
}
// with the hazard products (fs_params.hazard)
template <>
__global__ __launch_bounds__(256, FS_LTS_S2_WAVES) void sw_stage<1, true, true>(
    const DevInst *__restrict__ insts, const int4 *__restrict__ work, int nwork, Part *__restrict__ part, int m) {
  (void)nwork; (void)work;
  lts_stage_body<1, true>(insts, part, m);
}
// with flood barriers (extended faces)
template <>
__global__ __launch_bounds__(256, FS_LTS_S1_WAVES) void sw_stage<0, true, false, true>(
    const DevInst *__restrict__ insts, const int4 *__restrict__ work, int nwork, Part *__restrict__ part, int m) {
  (void)nwork; (void)work;
  lts_stage_body<0, false, true>(insts, part, m);
}
template <>
__global__ __launch_bounds__(256, FS_LTS_S2_WAVES) void sw_stage<1, true, false, true>(
    const DevInst *__restrict__ insts, const int4 *__restrict__ work, int nwork, Part *__restrict__ part, int m) {
  (void)nwork; (void)work;
  lts_stage_body<1, false, true>(insts, part, m);
}
template <>
__global__ __launch_bounds__(256, FS_LTS_S2_WAVES) void sw_stage<1, true, true, true>(
    const DevInst *__restrict__ insts, const int4 *__restrict__ work, int nwork, Part *__restrict__ part, int m) {
  (void)nwork; (void)work;
  lts_stage_body<1, true, true>(insts, part, m);
}

// Barrier books after each RK stage, one thread per barrier face (the barrier kernel,
// kernels.wgsl.js:366-412; solver-cpu.js:327-343): the theta-scaled flux the face passed this stage is
// booked while the barrier is up and water stands 2 cm over its crest; in the first stage the peak rate,
// the highest water and the hydrostatic load on the barrier, each with the step's start time (under local
// time stepping: a face of the level stepping now, timed at the cycle start). The evaluated state is U^n
// in the first stage and U* (the stepping level's cells) in the second.
template <int STAGE, bool LTS>
__global__ __launch_bounds__(64) void sw_barrier_books(const DevInst *__restrict__ insts, int inst, int m) {
  const DevInst &I = insts[inst];
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= I.nbar) return;
  if (STAGE == 1 && !I.secondOrder) return;
  m &= 0xff;
  const StepRec *rec = I.rec;
  const float dt = LTS ? rec->ldt[m] : rec->dt;
  if (!(dt > 0.f)) return;
  const double tnow = LTS ? rec->cycT0 : rec->t;
  const int4 bf = I.barFace[q];
  const int kb = bf.x, ka = bf.y;
  bool Ba = STAGE == 1, Bb = STAGE == 1;
  if constexpr (LTS) {
    const int la = tile_level(I, ka % I.pitch, ka / I.pitch), lb = tile_level(I, kb % I.pitch, kb / I.pitch);
    if (min(la, lb) != m) return;
    Ba = STAGE == 1 && la == m;
    Bb = STAGE == 1 && lb == m;
  }
  const float ha = Ba ? I.h1[ka] : I.h[ka], hb = Bb ? I.h1[kb] : I.h[kb];
  if (ha <= 0.f && hb <= 0.f) return;
  const float4 cr = I.crest[kb];
  const float crest = bf.z ? cr.z : cr.x;
  if (crest < -1e29f) return;
  const float ea = ha > 0.005f ? ha + I.z[ka] : -1e30f, eb = hb > 0.005f ? hb + I.z[kb] : -1e30f;
  const float w = fmaxf(ea, eb);
  const float f = I.barFlux[q];
  const bool over = w > crest + 0.02f;
  double *B = I.barBook + 9 * (size_t)q;
  if (over) {
    const double v = (double)f * (double)I.dx * (double)dt * (I.secondOrder ? 0.5 : 1.0);
    if (v > 0.0) B[0] += v;
    else B[1] -= v;
    if (B[8] < 0.0 && v != 0.0) B[8] = tnow;
  }
  if (STAGE == 0) {
    if (over && fabsf(f) > B[2]) { B[2] = fabsf(f); B[5] = tnow; }
    if (w > B[3]) { B[3] = w; B[6] = tnow; }
    const float gr = I.barGround[q];
    if (gr > -1e29f) {
      const float dA = fmaxf(0.f, fminf(ea, crest) - gr), dB = fmaxf(0.f, fminf(eb, crest) - gr);
      const double load = 0.5 * (double)I.g * fabs((double)dA * dA - (double)dB * dB);
      if (load > B[4]) { B[4] = load; B[7] = tnow; }
    }
  }
}

// ------------------------------------------------------------ local time stepping, per record

// After the stages of a level-m record, one kernel ends it: under first order a waiting cell's store takes the
// record's gathered flux (compensated, as kadd on U0 in ltsUpdate, kernels.wgsl.js:518; under second order the
// second stage has added it), then its first S blocks sum a slice of the partials each and the last does the
// books and the inlets. (A second stage that ended its record itself, each wavefront counting its slice done,
// measured 1.8x slower: a device-scope release after every wavefront writes back its XCD's L2.)
__global__ __launch_bounds__(256) void sw_lts_fold(const DevInst *__restrict__ insts, const Part *__restrict__ p1,
                                                   const Part *__restrict__ p2, int m) {
  const DevInst &I = insts[0];
  const int lane = threadIdx.x & 63;
  const int np = I.rec->lcount[m], G = lts_waves(I, m);
  const int w = xcd_wave_index_n(G, kLtsWaves);
  const bool two = I.secondOrder;
  if (w >= 0 && !two) {   // (second order: the second stage has added the gathers)
    for (int i = w; i < np; i += G) {
      const int4 wk = I.plist[m * I.capParts + i];
      const int X = wk.y - I.hl + lane;
      const bool mine = lane >= I.hl && lane < I.hl + I.own && X < I.nx;
      const int tx0 = wk.y / kLtsTile, ntx = min(I.own / kLtsTile, I.tx - tx0);
      const int yend = min(wk.z + wk.w, I.ny), tb0 = wk.z / kLtsTile, ntb = (yend - 1) / kLtsTile - tb0 + 1;
      // only a tile that waits beside a tile at m gathers: which tile rows of the part hold one (lane = tile
      // row * 8 + tile column; a part lies in one unit, at most 8 tile rows)
      const bool relt = (lane & 7) < ntx && (lane >> 3) < ntb && tile_waits(I, tx0 + (lane & 7), tb0 + (lane >> 3), m);
      const uint64_t rel = __ballot(relt);
      if (!rel || !mine) continue;
      // per tile row that holds one: the gather flags, then every flagged cell's store and gathers, all issued
      // before any is used, so the walk waits for memory twice per tile row rather than once per flagged row
      for (int yb = wk.z; yb < yend;) {
        const int ye = min(yend, (yb / kLtsTile + 1) * kLtsTile);
        if ((rel >> (8 * (yb / kLtsTile - tb0))) & 0xffull) {
          uint8_t f[kLtsTile];
          // (rows past the tile row's end load its last row's flag, then drop it: the loads need no branch)
#pragma unroll
          for (int q = 0; q < kLtsTile; ++q) f[q] = gld(I.gflag, min(yb + q, ye - 1) * I.pitch + X);
#pragma unroll
          for (int q = 0; q < kLtsTile; ++q) f[q] = yb + q < ye ? f[q] : (uint8_t)0;
            float sh[kLtsTile], sw[kLtsTile], shu[kLtsTile], shv[kLtsTile], a[kLtsTile], au[kLtsTile], av[kLtsTile];
#pragma unroll
          for (int q = 0; q < kLtsTile; ++q) {
            sh[q] = sw[q] = shu[q] = shv[q] = a[q] = au[q] = av[q] = 0.f;
            if (f[q]) {
              const int k = (yb + q) * I.pitch + X;
              const float4 t = gld(I.st4, k);
              sh[q] = t.x; sw[q] = t.y; shu[q] = t.z; shv[q] = t.w;
              const float4 ga = gld(I.g4, k);
              a[q] = ga.x; au[q] = ga.y; av[q] = ga.z;
            }
          }
#pragma unroll
          for (int q = 0; q < kLtsTile; ++q) {
            if (!f[q]) continue;
            const int k = (yb + q) * I.pitch + X;
            float h = sh[q], w = sw[q];
            kadd(h, w, a[q]);
            const float hu = shu[q] + au[q], hv = shv[q] + av[q];
            gst(I.st4, k, make_float4(h, w, hu, hv));
            gst(I.gflag, k, (uint8_t)0);
          }
        }
        yb = ye;
      }
    }
  }

  // ---- the record's end
  const float dt = I.rec->ldt[m];
  const int S = min(kLtsEndSlices, (int)gridDim.x), b = blockIdx.x;
  if (b >= S || !(dt > 0.f)) return;
  if (threadIdx.x < 64) {   // (one wavefront per block sums its slice)
    const int k0 = (int)((long long)G * b / S), k1 = (int)((long long)G * (b + 1) / S);
    if (lts_end_slice(I, p1, p2, b, k0, k1, S, lane)) lts_books(I, m, S, lane);
  }
}

// ------------------------------------------------------------ local time stepping, per cycle
// (_cycle, solver-cpu.js:504-527; ltsCycle/ltsLevel/ltsSmooth/ltsList, kernels.wgsl.js:677-741)

// max wave speed per 8x8 tile at the cycle start: one wavefront per 8 tiles of a tile row (lane = column, a
// walk down the tile row's 8 rows, then the maximum over each tile's 8 lanes), so every row's loads are
// contiguous; and the block's maximum into one of kSpeedSlots slots by atomic max: the cycle kernel reduces
// the slots (a max does not depend on the order of the atomics, and speeds are non-negative floats, which
// order as their bit patterns do)
__global__ __launch_bounds__(256) void sw_lts_speed(const DevInst *__restrict__ insts) {
  const DevInst &I = insts[0];
  const int t = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const int cbs = (I.tx + 7) / 8;   // blocks of 8 tiles across
  float s = 0.f;
  if (t < cbs * I.ty) {
    const int cb = t % cbs, tb = t / cbs;
    const int x = cb * 64 + lane, y0 = tb * kLtsTile;
    if (x < I.nx) {
      const int y1 = min(y0 + kLtsTile, I.ny);
      for (int y = y0; y < y1; ++y) {
        const int k = y * I.pitch + x;
        const float h = gld(I.h, k);
        const float hu = gld(I.hu, k), hv = gld(I.hv, k);
        if ((gld(I.mask, k) & MASK_OPEN) && h > I.hdry) s = fmaxf(s, cell_speed(h, hu, hv, I.g));
      }
    }
    s = fmaxf(s, __shfl_xor(s, 1, 64));
    s = fmaxf(s, __shfl_xor(s, 2, 64));
    s = fmaxf(s, __shfl_xor(s, 4, 64));
    const int tc = cb * 8 + (lane >> 3);
    if ((lane & 7) == 0 && tc < I.tx) I.tsp[tb * I.tx + tc] = s;
    s = wave_max(s);
  }
  __shared__ float s_m[4];
  if (lane == 0) s_m[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float b = fmaxf(fmaxf(s_m[0], s_m[1]), fmaxf(s_m[2], s_m[3]));
    if (b > 0.f) atomicMax(&I.speedSlots[blockIdx.x % kSpeedSlots], __float_as_uint(b));
  }
}

// cycle length D = min(2^L 0.9 cfl dx / smax (dtDry when dry), tEnd - t), base step D / 2^L
__global__ __launch_bounds__(1024) void sw_lts_cycle(const DevInst *__restrict__ insts) {
  const DevInst &I = insts[0];
  StepRec *rec = I.rec;
  __shared__ float s_m[16];
  float m = 0.f;
  // the tile speeds' maximum from sw_lts_speed's slots, which are cleared for the next cycle
  for (int t = threadIdx.x; t < kSpeedSlots; t += blockDim.x) {
    m = fmaxf(m, __uint_as_float(I.speedSlots[t]));
    I.speedSlots[t] = 0u;
  }
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int k = 1; k < (int)(blockDim.x >> 6); ++k) m = fmaxf(m, s_m[k]);
  const int P = 1 << I.L;
  const double t0 = rec->t, rem = rec->tEnd - t0;
  double D = 0.0, after = t0;
  if (rem > 0.0) {
    const double cand = m > 0.f ? (double)P * kLtsSafe * (double)I.cfl * (double)I.dx / (double)m : (double)I.dtDry;
    if (cand >= rem) { D = rem; after = rec->tEnd; }
    else { D = cand; after = t0 + cand; }
  }
  rec->smax = m;
  rec->cycT0 = t0;
  rec->cycD = D;
  rec->base = D / P;
  rec->tAfter = after;
  for (int l = 0; l <= kMaxLevels; ++l) rec->ldt[l] = l <= I.L ? (float)(rec->base * (double)(1 << l)) : 0.f;
}

// tile level: the largest m <= L whose step 2^m delta is within 0.9 of the CFL step of the fastest
// water in the tile and its 8 neighbours (all dry: L)
__global__ __launch_bounds__(256) void sw_lts_level(const DevInst *__restrict__ insts) {
  const DevInst &I = insts[0];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= I.tx * I.ty) return;
  const int x = t % I.tx, y = t / I.tx;
  float s = 0.f;
  for (int b = max(0, y - 1); b <= min(I.ty - 1, y + 1); ++b)
    for (int a = max(0, x - 1); a <= min(I.tx - 1, x + 1); ++a) s = fmaxf(s, I.tsp[b * I.tx + a]);
  const double dl = I.rec->base;
  int lv = I.L;
  if (s > 0.f && dl > 0.0) {
    const double q = floor(log2(kLtsSafe * (double)I.cfl * (double)I.dx / (double)s / dl));
    lv = (int)fmax(0.0, fmin((double)I.L, q));
  }
  I.lv0[t] = (uint8_t)lv;
}

// no tile more than one level above a face neighbour: the reference applies L sweeps of
// lv = min(lv, neighbour + 1); with levels in [0, L] that is exactly min over tiles within
// Manhattan distance L of (raw level + distance)
__global__ __launch_bounds__(256) void sw_lts_smooth(const DevInst *__restrict__ insts) {
  const DevInst &I = insts[0];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= I.tx * I.ty) return;
  const int x = t % I.tx, y = t / I.tx, L = I.L;
  int lv = I.lv0[t];
  for (int dy = -L; dy <= L; ++dy) {
    const int b = y + dy;
    if (b < 0 || b >= I.ty) continue;
    const int r = L - abs(dy);
    for (int dx = -r; dx <= r; ++dx) {
      const int a = x + dx;
      if (a < 0 || a >= I.tx) continue;
      lv = min(lv, (int)I.lv0[b * I.tx + a] + abs(dx) + abs(dy));
    }
  }
  I.lv[t] = (uint8_t)lv;
}

// per work unit (whole tiles): the rows a record of each level owns in it (the rule at lts_fill_gaps, above
// the LTS record work) and its own open cells per level
__global__ __launch_bounds__(256) void sw_lts_units(const DevInst *__restrict__ insts, const int4 *__restrict__ work) {
  const DevInst &I = insts[0];
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  const int nu = I.wend - I.wbeg;
  if (u >= nu) return;
  const int4 wk = work[I.wbeg + u];
  const int tx0 = wk.y / kLtsTile, tx1 = min(I.tx, (wk.y + I.own) / kLtsTile);
  const int ty0 = wk.z / kLtsTile, ty1 = min(I.ty, (wk.z + wk.w + kLtsTile - 1) / kLtsTile);
  int cells[kMaxLevels + 1] = {0, 0, 0, 0, 0};
  // per level, bit q = tile row q of the unit: every row owned, its first row owned, its last row owned
  unsigned full[kMaxLevels + 1] = {0, 0, 0, 0, 0}, first[kMaxLevels + 1] = {0, 0, 0, 0, 0};
  unsigned last[kMaxLevels + 1] = {0, 0, 0, 0, 0};
  const uint8_t *lv = I.lv;
  for (int b = ty0; b < ty1; ++b)
    for (int a = tx0; a < tx1; ++a) {
      const int t = b * I.tx + a, l = lv[t], op = I.tileOpen[t], bit = 1 << (b - ty0);
      cells[l] += op;
      if (!op) continue;
      full[l] |= bit;
      if (l > 0) {   // waits beside a tile at l - 1 in that level's records
        const int f = l - 1;
        if ((a > 0 && lv[t - 1] == f) || (a + 1 < I.tx && lv[t + 1] == f)) full[f] |= bit;
        if (b > 0 && lv[t - I.tx] == f) first[f] |= bit;
        if (b + 1 < I.ty && lv[t + I.tx] == f) last[f] |= bit;
      }
    }
  const int nr = min(wk.w, I.ny - wk.z);   // rows of the unit (at most 64)
  const uint64_t in = nr >= 64 ? ~0ull : (1ull << nr) - 1ull;
  for (int l = 0; l <= kMaxLevels; ++l) {
    uint64_t msk = 0;
    for (int q = 0; q < ty1 - ty0; ++q) {
      if ((full[l] >> q) & 1u) msk |= 0xffull << (kLtsTile * q);
      if ((first[l] >> q) & 1u) msk |= 1ull << (kLtsTile * q);
      if ((last[l] >> q) & 1u) msk |= 1ull << (kLtsTile * q + kLtsTile - 1);
    }
    I.urows[(size_t)l * nu + u] = l <= I.L ? lts_fill_gaps(msk & in) : 0ull;
    I.ucells[u * (kMaxLevels + 1) + l] = cells[l];
  }
}

// parts of a unit's row mask for each part size (candidates 1, 2, 4, ... 32 rows, then whole runs)
__device__ __forceinline__ void lts_count(uint64_t msk, int cnt[kLtsCand]) {
  lts_runs(msk, [&](int s, int len) {
    (void)s;
#pragma unroll
    for (int c = 0; c < kLtsCand - 1; ++c) cnt[c] += (len + (1 << c) - 1) >> c;
    cnt[kLtsCand - 1] += 1;
  });
}

// per level (one block each): the parts of the level's records in unit order (a block-wide ordered scan:
// deterministic), with the part size that keeps a record of few owned rows within kLtsFineWaves; the own open
// cells, and (rain field) the rain multiplier summed over them
__global__ __launch_bounds__(1024) void sw_lts_lists(const DevInst *__restrict__ insts, const int4 *__restrict__ work) {
  const DevInst &I = insts[0];
  const int l = blockIdx.x, nu = I.wend - I.wbeg, T = blockDim.x, tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6, nwv = T >> 6;
  __shared__ double s_red[16][kLtsCand + 1];
  __shared__ int s_off[16];
  __shared__ int s_cand;
  if (l > I.L) {
    if (tid == 0) { I.rec->lcount[l] = 0; I.rec->lcells[l] = 0; }
    return;
  }
  const uint64_t *um = I.urows + (size_t)l * nu;
  // totals of the part counts of each candidate size, and of the level's own open cells
  double acc[kLtsCand + 1];
#pragma unroll
  for (int c = 0; c <= kLtsCand; ++c) acc[c] = 0.0;
  for (int u = tid; u < nu; u += T) {
    int cnt[kLtsCand] = {0, 0, 0, 0, 0, 0, 0};
    lts_count(um[u], cnt);
#pragma unroll
    for (int c = 0; c < kLtsCand; ++c) acc[c] += (double)cnt[c];
    acc[kLtsCand] += (double)I.ucells[u * (kMaxLevels + 1) + l];
  }
#pragma unroll
  for (int c = 0; c <= kLtsCand; ++c) {
    const double v = wave_sum(acc[c]);
    if (lane == 0) s_red[wv][c] = v;
  }
  __syncthreads();
  if (tid == 0) {
    double tot[kLtsCand + 1];
    for (int c = 0; c <= kLtsCand; ++c) {
      tot[c] = 0.0;
      for (int k = 0; k < nwv; ++k) tot[c] += s_red[k][c];
    }
    const double cap = (double)min(kLtsFineWaves, I.capParts);
    int cand = kLtsBigCand;
    for (int c = 0; c < kLtsBigCand; ++c)
      if (tot[c] <= cap) { cand = c; break; }
    s_cand = cand;
    I.rec->lcount[l] = (int)tot[cand];
    I.rec->lcells[l] = (int)tot[kLtsCand];
  }
  __syncthreads();
  const int rp = lts_cand_rows(s_cand);
  int4 *out = I.plist + (size_t)l * I.capParts;
  int base = 0;
  for (int u0 = 0; u0 < nu; u0 += T) {   // chunks of T units, scanned in unit order
    const int u = u0 + tid;
    const uint64_t msk = u < nu ? um[u] : 0ull;
    int n = 0;
    lts_runs(msk, [&](int s, int len) { (void)s; n += (len + rp - 1) / rp; });
    int x = n;   // inclusive scan over the wave
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) s_off[wv] = x;
    __syncthreads();
    int before = base, chunk = 0;
    for (int k = 0; k < nwv; ++k) {
      if (k < wv) before += s_off[k];
      chunk += s_off[k];
    }
    __syncthreads();   // (s_off is rewritten by the next chunk)
    if (n) {
      const int4 wk = work[I.wbeg + u];
      int pos = before + x - n;
      lts_runs(msk, [&](int s, int len) {
        for (int r = 0; r < len; r += rp) out[pos++] = make_int4(wk.x, wk.y, wk.z + s + r, min(rp, len - r));
      });
    }
    base += chunk;
  }
  if (I.tileRf) {   // rain multiplier of the level (only with a rain field): fixed-order sums
    double a = 0.0;
    for (int t = tid; t < I.tx * I.ty; t += T)
      if (I.lv[t] == l) a += (double)I.tileRf[t];
    a = wave_sum(a);
    if (lane == 0) s_red[wv][0] = a;
    __syncthreads();
    if (tid == 0) {
      double r = 0.0;
      for (int k = 0; k < nwv; ++k) r += s_red[k][0];
      I.rec->lrf[l] = r;
    }
  }
}

// end of a cycle: the clock (solver-cpu.js:525)
__global__ void sw_lts_cycle_end(const DevInst *__restrict__ insts) {
  StepRec *rec = insts[0].rec;
  if (rec->cycD > 0.0) {
    rec->t = rec->tAfter;
    rec->steps += 1;
  }
}

// max wave speed |u| + |v| + 2 sqrt(g h) over wet open cells (dtReduceAt, kernels.wgsl.js:767-780);
// used before the first step and after the host changes the state
__global__ __launch_bounds__(256) void sw_speed(const DevInst *__restrict__ insts, const int4 *__restrict__ work,
                                                int nwork, Part *__restrict__ part) {
  const int wid = xcd_wave_index(4);
  if (wid >= nwork) return;
  const int lane = threadIdx.x & 63;
  const int4 wk = work[wid];
  const DevInst &I = insts[wk.x];
  const int X = wk.y - I.hl + lane;
  float smax = 0.f;
  if (lane >= I.hl && lane < I.hl + I.own && X < I.nx) {
    for (int y = wk.z; y < min(wk.z + wk.w, I.ny); ++y) {
      const size_t k = (size_t)y * I.pitch + X;
      const float h = I.h[k];
      if ((I.mask[k] & MASK_OPEN) && h > I.hdry) smax = fmaxf(smax, cell_speed(h, I.hu[k], I.hv[k], I.g));
    }
  }
  const float s = wave_max(smax);
  if (lane == 0) part[wid] = Part{0.0, 0.0, 0.0, s, 0};
}

// first order (secondOrder = false): the Euler stage's U* plus the sources becomes U^{n+1}
// (step, solver-cpu.js:432-438: one stage, then _sources). LTS: the level-m cells of a record,
// whose store is then spent (solver-cpu.js:550-557).
template <bool LTS, bool HZ, bool EXT = false>
__device__ __forceinline__ bool euler_finish_strip(const DevInst &I, const int4 wk, int lane, int m, float dt, Part *part,
                                                   int cells, float tnow, bool acc = false) {
  double acc_inf = 0.0;
  float smax = 0.f;
  bool stir = false;
  const int X = wk.y - I.hl + lane;
  if (dt > 0.f && lane >= I.hl && lane < I.hl + I.own && X < I.nx) {
    const float rain = I.rec->rainRate, areaf = I.dx * I.dx;
    for (int y = wk.z; y < min(wk.z + wk.w, I.ny); ++y) {
      const int k = y * I.pitch + X;
      if (!(gld(I.mask, k) & MASK_OPEN)) continue;
      if (LTS && tile_level(I, X, y) != m) continue;
      float qh = gld(I.h1, k), qw = gld(I.w1, k), qhu = gld(I.hu1, k), qhv = gld(I.hv1, k);
      sources_cell<EXT, EXT || LTS, !LTS>(I, k, dt, rain, areaf, qh, qw, qhu, qhv, acc_inf, smax);
      if constexpr (HZ) {
        if (I.hzA) hazard_cell(I, k, qh, qhu, qhv, dt, tnow);
      }
      gst(I.h, k, qh); gst(I.w, k, qw); gst(I.hu, k, qhu); gst(I.hv, k, qhv);
      if (LTS) gst(I.st4, k, make_float4(0.f, 0.f, 0.f, 0.f));
      stir = stir || qh != 0.f || qw != 0.f || qhu != 0.f || qhv != 0.f;
    }
  }
  const double f = wave_sum(acc_inf);
  const float s = wave_max(smax);
  if (lane == 0) {
    if (acc) {   // a further part of the same wavefront (LTS)
      Part q = *part;
      q.infil += f; q.smax = fmaxf(q.smax, s); q.cells += cells;
      *part = q;
    } else {
      *part = Part{0.0, 0.0, f, s, cells};
    }
  }
  return __any(stir);
}

template <bool LTS, bool HZ = false, bool EXT = false>
__global__ __launch_bounds__(256) void sw_euler_finish(const DevInst *__restrict__ insts,
                                                       const int4 *__restrict__ work, int nwork,
                                                       Part *__restrict__ part, int m) {
  const int wid = xcd_wave_index(kGlobalWaves);
  const int lane = threadIdx.x & 63;
  if constexpr (!LTS) {
    if (wid >= nwork) return;
    const int4 wk = work[wid];
    const DevInst &I = insts[wk.x];
    if (I.secondOrder) return;  // second order: sw_stage<1> writes this strip's partial
    if ((I.hzA != nullptr) != HZ || ext_faces(I) != EXT) return;  // another variant advances it
    const int u = wid - I.wbeg;
    if (I.wet && !I.act[u]) {
      if (lane == 0) part[wid] = Part{0.0, 0.0, 0.0, 0.f, 0};
      return;
    }
    const float dt = I.rec->dt;
    const bool stir = euler_finish_strip<false, HZ, EXT>(I, wk, lane, 0, dt, part + wid, dt > 0.f ? I.sopen[u] : 0,
                                                     (float)I.rec->tAfter);
    if (I.wet && dt > 0.f && lane == 0) I.wet[u] = stir ? 1 : 0;
  } else {
    const DevInst &I = insts[0];
    if (I.secondOrder) return;
    const int kk = m >> 8;
    m &= 0xff;
    const float dt = I.rec->ldt[m];
    const int np = I.rec->lcount[m], G = lts_waves(I, m);
    const int w = xcd_wave_index_n(G, kLtsWaves);
    if (w < 0) return;
    bool acc = false;
    for (int i = w; i < np; i += G) {
      euler_finish_strip<true, HZ, EXT>(I, I.plist[m * I.capParts + i], lane, m, dt, part + w, 0,
                                        lts_record_end_time(I, kk), acc);
      acc = true;
    }
    if (!acc && lane == 0) part[w] = Part{0.0, 0.0, 0.0, 0.f, 0};
  }
}

// inlet exchange after the sources (exchangeAt, kernels.wgsl.js:597-618): the held rate over the
// step, never more water than the cell holds; negative rates are point inflows. Writes the wave
// speed of the inlet cell for the next dt (the second stage leaves inlet cells out of its maximum).
// LTS: inlets whose cell is at the record's level, with its step (exchangeL, kernels.wgsl.js:596).
template <bool LTS>
__device__ __forceinline__ void exchange_inlet(const DevInst &I, const int j, const int m) {
  Inlet &q = I.inl[j];
  float dt = I.rec->dt;
  const size_t k = (size_t)q.cell;
  if constexpr (LTS) {
    if (tile_level(I, (int)(k % I.pitch), (int)(k / I.pitch)) != m) return;
    dt = I.rec->ldt[m];
  }
  float h = I.h[k], w = I.w[k], hu = I.hu[k], hv = I.hv[k];
  if (dt > 0.f) {
    const float A = I.dx * I.dx;
    const float dd = fminf(q.q * dt / A, fmaxf(h, 0.f));
    kadd(h, w, -dd);
    if (h <= 0.f) { h = 0.f; w = 0.f; }
    if (h <= I.hdry) { hu = 0.f; hv = 0.f; }
    I.h[k] = h; I.w[k] = w; I.hu[k] = hu; I.hv[k] = hv;
    q.vol += (double)dd * (double)A;
    // water put into a quiet strip wakes it for the next step
    if (!LTS && I.wet && (h != 0.f || w != 0.f)) {
      const int x = (int)(k % I.pitch), y = (int)(k / I.pitch);
      I.wet[(y / I.rows) * I.sxn + x / I.own] = 1;
    }
  }
  q.h = h;
  I.ispeed[j] = h > I.hdry ? cell_speed(h, hu, hv, I.g) : 0.f;
  // the water arriving at the inlet cell, for a grate law that sees moving water (approachAt,
  // kernels.wgsl.js:620-637): each interior neighbour's unit discharge toward the cell; the largest is the
  // frontal flow, the rest arrives from the side. Reported only: it changes no water.
  const int x = (int)(k % I.pitch), y = (int)(k / I.pitch);
  const float fx = x > 0 ? fmaxf(0.f, I.hu[k - 1]) : 0.f;
  const float fy = x + 1 < I.nx ? fmaxf(0.f, -I.hu[k + 1]) : 0.f;
  const float fz = y > 0 ? fmaxf(0.f, I.hv[k - I.pitch]) : 0.f;
  const float fw = y + 1 < I.ny ? fmaxf(0.f, -I.hv[k + I.pitch]) : 0.f;
  float fm = fx, ax = 0.f;
  size_t nb = k - 1;
  if (fy > fm) { fm = fy; nb = k + 1; }
  if (fz > fm) { fm = fz; nb = k - I.pitch; ax = 1.f; }
  if (fw > fm) { fm = fw; nb = k + I.pitch; ax = 1.f; }
  const float hn = fm > 0.f ? I.h[nb] : 0.f;
  I.iapp[j] = make_float4(fm * I.dx, (fx + fy + fz + fw - fm) * I.dx, hn, ax);
}

template <bool LTS>
__global__ void sw_exchange(const DevInst *__restrict__ insts, int inst, int m) {
  const DevInst &I = insts[inst];
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < I.nin) exchange_inlet<LTS>(I, j, m);
}

// End of a step, one block per instance (dtFinal, kernels.wgsl.js:783-799): fold the stage
// partials into the FP64 books in a fixed order, advance the clock, and set the next dt from the
// max wave speed: dt = cfl dx / smax (dtDry when dry), clamped to land exactly on tEnd.
// mode 0: end of a step; 1: start of an advance: instances the host marked dirty take their max
// speed from the sw_speed partials (p2), the others re-clamp their last one to the new tEnd.
__global__ __launch_bounds__(1024) void sw_dt_final(const DevInst *__restrict__ insts, const Part *__restrict__ p1,
                                                   const Part *__restrict__ p2, int mode) {
  const DevInst &I = insts[blockIdx.x];
  StepRec *rec = I.rec;
  const int t = threadIdx.x;
  __shared__ double s_o[16], s_c[16], s_f[16];
  __shared__ float s_m[16];
  __shared__ int s_n[16], s_s[16];
  double o = 0.0, c = 0.0, f = 0.0;
  float m = 0.f;
  int cells = 0, strips = 0;
  const bool fresh = mode == 0 || rec->dirty;
  if (fresh) {
#pragma unroll 8
    for (int k = I.wbeg + t; k < I.wend; k += blockDim.x) {
      const Part b = p2[k];
      m = fmaxf(m, b.smax);
      if (mode == 0) {
        const Part a = p1[k];
        o += a.out + b.out;
        c += a.clip + b.clip;
        f += b.infil;
        cells += b.cells;
        if (I.wet) strips += I.act[k - I.wbeg] ? 1 : 0;
      }
    }
    if (mode == 0)
      for (int k = t; k < I.nin; k += blockDim.x) m = fmaxf(m, I.ispeed[k]);
  }
  o = wave_sum(o); c = wave_sum(c); f = wave_sum(f); m = wave_max(m);
  cells = wave_sum_i(cells);
  strips = wave_sum_i(strips);
  if ((t & 63) == 0) {
    s_o[t >> 6] = o; s_c[t >> 6] = c; s_f[t >> 6] = f; s_m[t >> 6] = m; s_n[t >> 6] = cells; s_s[t >> 6] = strips;
  }
  __syncthreads();
  if (t != 0) return;
  for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
    o += s_o[k]; c += s_c[k]; f += s_f[k]; m = fmaxf(m, s_m[k]); cells += s_n[k]; strips += s_s[k];
  }
  // a no-op step (dt = 0, past tEnd) measured nothing: keep the last real max speed
  if (!fresh || (mode == 0 && !(rec->dt > 0.f))) m = rec->smax;
  rec->dirty = 0;
  if (mode == 0 && rec->dt > 0.f) {
    const float dtp = rec->dt;
    rec->outflow += o;
    rec->clipped += c;
    rec->infil += f;
    if (rec->rainCells >= 0) rec->rain += (double)(rec->rainRate * dtp) * (double)rec->rainCells * (double)I.area;
    else rec->rain += (double)rec->rainRate * (double)dtp * rec->rainArea;
    rec->t = rec->tAfter;
    rec->steps += 1;
    rec->lastDt = dtp;
    rec->cellUpdates += (double)cells;
    rec->active = cells;
    rec->activeStrips = strips;
  }
  rec->smax = m;
  const double rem = rec->tEnd - rec->t;
  float dt = 0.f;
  double after = rec->t;
  if (rem > 0.0) {
    const float cand = m > 0.f ? I.cfl * I.dx / m : I.dtDry;
    if ((double)cand >= rem) { dt = (float)rem; after = rec->tEnd; }
    else { dt = cand; after = rec->t + (double)cand; }
    if (!(dt > 0.f)) { dt = 0.f; after = rec->t; }
  }
  rec->dt = dt;
  rec->tAfter = after;
}

}  // namespace fs

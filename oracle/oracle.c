/*
 * FP64 CPU oracle for the 2D shallow-water step. TEST INFRASTRUCTURE ONLY:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, and only as the checker. The product (floodsim_amd) never
 * links or calls it.
 *
 * A restatement in C of the reference CPU solver, src/solver-cpu.js, in plain
 * FP64 with the same operation order (built with -ffp-contract=off, so no
 * fused multiply-add changes a rounding):
 *   ghost fill        _fillGhostStatic / _fillGhost   solver-cpu.js:165-203
 *   primitives        _prims                           solver-cpu.js:207-213
 *   MUSCL             _recon, minmod, vanleer          solver-cpu.js:17-24, 216-233
 *   hydrostatic face  _face                            solver-cpu.js:237-267
 *   HLLC              hllc                             solver-cpu.js:579-602
 *   operator L(U)     _L (theta limiter, dU, qout)     solver-cpu.js:291-381
 *   time step         computeDt, step, _fix            solver-cpu.js:383-441
 *   sources           _sources (friction, rain, inf.)  solver-cpu.js:456-488
 *   local stepping    _cycle, _record                  solver-cpu.js:504-558
 *   books             volume, massBalance              solver-cpu.js:156-162, 564-568
 *   inlet exchange    CpuSim.advance                   benchmarks/physics/street/exec-cpu.mjs:46-55
 *                     (per LTS record as the GPU's exchangeL, kernels.wgsl.js:596)
 *   F26 options       implicitFriction (_friction, the implicit factor of _sources) and stepFlux
 *                     (the cascade clause of _recon, _face's step flux, _stepFlux)
 *                                                      solver-cpu.js:222-225, 250-273, 419-423, 443-452, 465-467, 542-545
 *   water-level edge  stageCells, setStage, the stage ghosts of _fillGhost
 *                                                      solver-cpu.js:109-117, 197-204
 *   flood barriers    crest faces, _weir, setCrestLevels, the barrier books of _L
 *                                                      solver-cpu.js:98-105, 136-139, 244-258, 277-287, 327-343
 *   Green-Ampt, GI    the ga / gi options and their terms of _sources  solver-cpu.js:94-96, 466-474
 *
 * Loops over rows may run on several OpenMP threads. Every per-cell update is
 * independent; sums (rain, infiltration, clipping) are taken per row and then
 * added in row order, so results do not depend on the thread count.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define GH 2
#define LTS_TILE 8
#define LTS_SAFE 0.9

enum { BC_WALL = 0, BC_OPEN = 1, BC_OUTFLOW = 2 };

typedef struct {
  int nx, ny;
  double dx, g, cfl, hdry, hfront, dtDry;
  int secondOrder, limiter; /* limiter 0 minmod, 1 van Leer */
  int bc[4];                /* west, east, south, north */
  int lts;                  /* 0 = global stepping, else levels 1..4 */
  int implicitFriction;     /* F26: friction implicit in the new speed, and on the predictor the second stage sees */
  int stepFlux;             /* F26: a film stepping down a bed step taller than itself passes its own discharge */
} or_params;

typedef struct {
  or_params p;
  int NX, NY, N;
  double *h, *hu, *hv, *z, *n, *infil, *rainField;
  uint8_t *M;
  double rainRate;
  double *Hp, *Ep, *Up, *Vp, *h0, *hu0, *hv0, *dh, *dhu, *dhv, *FX, *FY, *maxDepth, *TH;
  double *keepU, *keepV; /* implicitFriction: the predictor's momentum, restored after the second L */
  /* water-level boundary: [ghost, edge cell] pairs of the stage cells, and the stage (absolute, m) */
  int nStage;
  int *stageG, *stageS;
  double stageLevel;
  /* flood barriers: crest (absolute, -INFINITY = none) and weir coefficient per x-face (at the right cell) and
   * y-face (at the upper cell); the barrier faces (upper/right cell, axis, ground) and their books, 9 per
   * face: plus, minus, peak, level, load, tPeak, tLevel, tLoad, tOver */
  double *CX, *CY, *WX, *WY;
  int nBar;
  int *barC, *barAxis;
  double *barGround, *bar;
  double dtCur; /* the step opL evaluates (the weir's flow cap) */
  /* Green-Ampt soil (psiD, Fmax, pervious share; F infiltrated so far) and GI stores (rate, capacity, drain;
   * stored depth), per padded cell, zero where absent */
  double *gaPsiD, *gaFmax, *gaPerv, *gaF, *giRate, *giCap, *giDrain, *giStore;
  double *rowA, *rowB; /* per-row partial sums */
  double t;
  long steps;
  double volRain, volInfil, volOutflow, volClipped, volInitial;
  /* local time stepping */
  int L, tilesX, tilesY, lev, stage;
  int *LV, *TI;
  double *tsp, *aH, *aHu, *aHv;
  long cellSteps, records;
  /* inlets */
  int nIn;
  int *inC;
  double *inQ, *inV;
  double exchanged;
  /* hazard products per interior cell (HazardTracker, hazard.js:75-96): peak d*v, peak force per
   * width, time deeper than 15 cm and when first, peak speed and when first deeper than VCUT, and the
   * worst class per standard */
  double *hzDv, *hzForce, *hzDur, *hzArr, *hzVmax, *hzWet;
  uint8_t *hzCls; /* 3 per cell: ARR H1-H6, FD2320, FD2320 urban */
} or_solver;

static double minmod(double a, double b) {
  if (a * b <= 0) return 0;
  return fabs(a) < fabs(b) ? a : b;
}
static double vanleer(double a, double b) {
  double p = a * b;
  return p <= 0 ? 0 : (2 * p) / (a + b);
}
static inline double lim(const or_solver *s, double a, double b) { return s->p.limiter ? vanleer(a, b) : minmod(a, b); }
static inline double dmax(double a, double b) { return a > b ? a : (a == a ? b : a); }
static inline double dmin(double a, double b) { return a < b ? a : (a == a ? b : a); }
static inline int lvc(const or_solver *s, int c) { return s->LV[s->TI[c]]; }

static double *zalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

void or_destroy(or_solver *s);

or_solver *or_create(const or_params *p, const double *z, const double *manning, double manningScalar,
                     const uint8_t *mask) {
  or_solver *s = (or_solver *)calloc(1, sizeof(or_solver));
  s->p = *p;
  s->NX = p->nx + 2 * GH;
  s->NY = p->ny + 2 * GH;
  s->N = s->NX * s->NY;
  size_t N = (size_t)s->N;
  s->h = zalloc(N); s->hu = zalloc(N); s->hv = zalloc(N);
  s->z = zalloc(N); s->n = zalloc(N); s->M = (uint8_t *)calloc(N, 1);
  s->infil = zalloc(N); s->rainField = NULL;
  s->Hp = zalloc(N); s->Ep = zalloc(N); s->Up = zalloc(N); s->Vp = zalloc(N);
  s->h0 = zalloc(N); s->hu0 = zalloc(N); s->hv0 = zalloc(N);
  s->dh = zalloc(N); s->dhu = zalloc(N); s->dhv = zalloc(N);
  s->FX = zalloc(N * 9); s->FY = zalloc(N * 9);
  s->maxDepth = zalloc(N); s->TH = zalloc(N);
  s->keepU = zalloc(N); s->keepV = zalloc(N);
  s->rowA = zalloc((size_t)s->NY); s->rowB = zalloc((size_t)s->NY);
  s->lev = -1;
  s->stageLevel = -1e9;
  s->CX = zalloc(N); s->CY = zalloc(N); s->WX = zalloc(N); s->WY = zalloc(N);
  s->gaPsiD = zalloc(N); s->gaFmax = zalloc(N); s->gaPerv = zalloc(N); s->gaF = zalloc(N);
  s->giRate = zalloc(N); s->giCap = zalloc(N); s->giDrain = zalloc(N); s->giStore = zalloc(N);
  for (size_t c = 0; c < N; c++) { s->CX[c] = -INFINITY; s->CY[c] = -INFINITY; }
  {
    size_t n = (size_t)p->nx * p->ny;
    s->hzDv = zalloc(n); s->hzForce = zalloc(n); s->hzDur = zalloc(n); s->hzArr = zalloc(n);
    s->hzVmax = zalloc(n); s->hzWet = zalloc(n); s->hzCls = (uint8_t *)calloc(3 * n, 1);
    for (size_t k = 0; k < n; k++) { s->hzArr[k] = NAN; s->hzWet[k] = NAN; }
  }
  for (int j = 0; j < p->ny; j++)
    for (int i = 0; i < p->nx; i++) {
      int c = (j + GH) * s->NX + i + GH, k = j * p->nx + i;
      s->z[c] = z[k];
      s->n[c] = manning ? manning[k] : manningScalar;
      s->M[c] = mask ? mask[k] : 1;
    }
  /* _fillGhostStatic: bed, roughness and mask into ghosts, zero gradient */
  for (int j = 0; j < s->NY; j++)
    for (int i = 0; i < s->NX; i++) {
      if (i >= GH && i < s->NX - GH && j >= GH && j < s->NY - GH) continue;
      int ii = i < GH ? GH : (i > s->NX - GH - 1 ? s->NX - GH - 1 : i);
      int jj = j < GH ? GH : (j > s->NY - GH - 1 ? s->NY - GH - 1 : j);
      int c = j * s->NX + i, src = jj * s->NX + ii;
      s->z[c] = s->z[src]; s->n[c] = s->n[src]; s->M[c] = 1;
    }
  if (p->lts) {
    s->L = p->lts < 1 ? 1 : (p->lts > 4 ? 4 : p->lts);
    int tx = (p->nx + LTS_TILE - 1) / LTS_TILE, ty = (p->ny + LTS_TILE - 1) / LTS_TILE;
    s->tilesX = tx; s->tilesY = ty;
    s->LV = (int *)calloc((size_t)tx * ty, sizeof(int));
    s->TI = (int *)calloc(N, sizeof(int));
    s->tsp = zalloc((size_t)tx * ty);
    for (int j = 0; j < s->NY; j++)
      for (int i = 0; i < s->NX; i++) {
        int ii = i - GH < 0 ? 0 : (i - GH > p->nx - 1 ? p->nx - 1 : i - GH);
        int jj = j - GH < 0 ? 0 : (j - GH > p->ny - 1 ? p->ny - 1 : j - GH);
        s->TI[j * s->NX + i] = (jj / LTS_TILE) * tx + ii / LTS_TILE;
      }
    s->aH = zalloc(N); s->aHu = zalloc(N); s->aHv = zalloc(N);
  }
  return s;
}

void or_destroy(or_solver *s) {
  if (!s) return;
  double *arrs[] = {s->h, s->hu, s->hv, s->z, s->n, s->infil, s->rainField, s->Hp, s->Ep, s->Up, s->Vp,
                    s->h0, s->hu0, s->hv0, s->dh, s->dhu, s->dhv, s->FX, s->FY, s->maxDepth, s->TH,
                    s->rowA, s->rowB, s->tsp, s->aH, s->aHu, s->aHv, s->inQ, s->inV, s->keepU, s->keepV};
  for (size_t k = 0; k < sizeof(arrs) / sizeof(arrs[0]); k++) free(arrs[k]);
  free(s->M); free(s->LV); free(s->TI); free(s->inC); free(s->stageG); free(s->stageS);
  free(s->CX); free(s->CY); free(s->WX); free(s->WY); free(s->barC); free(s->barAxis); free(s->barGround); free(s->bar);
  free(s->gaPsiD); free(s->gaFmax); free(s->gaPerv); free(s->gaF);
  free(s->giRate); free(s->giCap); free(s->giDrain); free(s->giStore);
  free(s->hzDv); free(s->hzForce); free(s->hzDur); free(s->hzArr); free(s->hzVmax); free(s->hzWet); free(s->hzCls);
  free(s);
}

void or_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

/* interior arrays (nx*ny, row-major) */
void or_set_infil(or_solver *s, const double *infil) {
  for (int j = 0; j < s->p.ny; j++)
    for (int i = 0; i < s->p.nx; i++) s->infil[(j + GH) * s->NX + i + GH] = infil[j * s->p.nx + i];
}
void or_set_rain_field(or_solver *s, const double *rf) {
  if (!s->rainField) {
    s->rainField = zalloc((size_t)s->N);
    for (int c = 0; c < s->N; c++) s->rainField[c] = 1;
  }
  for (int j = 0; j < s->p.ny; j++)
    for (int i = 0; i < s->p.nx; i++) s->rainField[(j + GH) * s->NX + i + GH] = rf[j * s->p.nx + i];
}
void or_set_rain_rate(or_solver *s, double r) { s->rainRate = r; }

double or_volume(const or_solver *s) {
  double v = 0;
  for (int j = 0; j < s->p.ny; j++)
    for (int i = 0; i < s->p.nx; i++) {
      int c = (j + GH) * s->NX + i + GH;
      if (s->M[c]) v += s->h[c];
    }
  return v * s->p.dx * s->p.dx;
}

/* setDepth (solver-cpu.js:142-148): h = mask ? max(0, value) : 0; books the initial volume */
void or_set_depth(or_solver *s, const double *h) {
  for (int j = 0; j < s->p.ny; j++)
    for (int i = 0; i < s->p.nx; i++) {
      int c = (j + GH) * s->NX + i + GH;
      double v = h[j * s->p.nx + i];
      s->h[c] = s->M[c] ? dmax(0, v) : 0;
    }
  s->volInitial = or_volume(s);
}
/* setVelocity (solver-cpu.js:149-154): hu = h*u, hv = h*v */
void or_set_velocity(or_solver *s, const double *u, const double *v) {
  for (int j = 0; j < s->p.ny; j++)
    for (int i = 0; i < s->p.nx; i++) {
      int c = (j + GH) * s->NX + i + GH, k = j * s->p.nx + i;
      s->hu[c] = s->h[c] * u[k];
      s->hv[c] = s->h[c] * v[k];
    }
}
/* raw state (for windows restarted from a saved state): no book is touched */
void or_set_state(or_solver *s, const double *h, const double *hu, const double *hv) {
  for (int j = 0; j < s->p.ny; j++)
    for (int i = 0; i < s->p.nx; i++) {
      int c = (j + GH) * s->NX + i + GH, k = j * s->p.nx + i;
      s->h[c] = h[k]; s->hu[c] = hu[k]; s->hv[c] = hv[k];
    }
}
void or_set_inlets(or_solver *s, const int *cells, int n) {
  free(s->inC); free(s->inQ); free(s->inV);
  s->nIn = n;
  s->inC = (int *)calloc(n ? n : 1, sizeof(int));
  s->inQ = zalloc(n); s->inV = zalloc(n);
  for (int m = 0; m < n; m++) {
    int k = cells[m], i = k % s->p.nx, j = k / s->p.nx;
    s->inC[m] = (j + GH) * s->NX + i + GH;
  }
}
void or_set_inlet_rates(or_solver *s, const double *q) {
  for (int m = 0; m < s->nIn; m++) s->inQ[m] = q[m];
}
/* readInlets: volume since the last call (m^3) and depth; resets the volumes */
void or_read_inlets(or_solver *s, double *vol, double *h) {
  for (int m = 0; m < s->nIn; m++) {
    if (vol) vol[m] = s->inV[m];
    if (h) h[m] = s->h[s->inC[m]];
    s->inV[m] = 0;
  }
}

/* ---------- ghost cells (solver-cpu.js:174-203) ---------- */
static double outVel(const or_solver *s, const double *h, int e, int i) {
  double he = h[e];
  if (he <= s->p.hdry || h[i] <= s->p.hdry || !s->M[i]) return 0;
  double sl = ((h[i] + s->z[i]) - (he + s->z[e])) / s->p.dx - 1e-4;
  if (sl <= 0) return 0;
  double crit = sqrt(s->p.g * he);
  return s->n[e] > 0 ? dmin(crit, pow(he, 2.0 / 3.0) * sqrt(sl) / s->n[e]) : crit;
}
static void fillGhost(or_solver *s, double *h, double *hu, double *hv) {
  const int NX = s->NX, NY = s->NY;
  const int *bc = s->p.bc;
#define PUT(c, src, su, sv) do { h[c] = h[src]; hu[c] = (su) * hu[src]; hv[c] = (sv) * hv[src]; } while (0)
  for (int j = GH; j < NY - GH; j++)
    for (int g = 0; g < GH; g++) {
      int c = j * NX + (GH - 1 - g), src = j * NX + (bc[0] == BC_WALL ? GH + g : GH);
      PUT(c, src, bc[0] == BC_WALL ? -1.0 : 1.0, 1.0);
      if (bc[0] == BC_OUTFLOW) hu[c] = -h[c] * outVel(s, h, src, src + 1);
      c = j * NX + (NX - GH + g);
      src = j * NX + (bc[1] == BC_WALL ? NX - GH - 1 - g : NX - GH - 1);
      PUT(c, src, bc[1] == BC_WALL ? -1.0 : 1.0, 1.0);
      if (bc[1] == BC_OUTFLOW) hu[c] = h[c] * outVel(s, h, src, src - 1);
    }
  for (int i = 0; i < NX; i++)
    for (int g = 0; g < GH; g++) {
      int c = (GH - 1 - g) * NX + i, src = (bc[2] == BC_WALL ? GH + g : GH) * NX + i;
      PUT(c, src, 1.0, bc[2] == BC_WALL ? -1.0 : 1.0);
      if (bc[2] == BC_OUTFLOW) hv[c] = -h[c] * outVel(s, h, src, src + NX);
      c = (NY - GH + g) * NX + i;
      src = (bc[3] == BC_WALL ? NY - GH - 1 - g : NY - GH - 1) * NX + i;
      PUT(c, src, 1.0, bc[3] == BC_WALL ? -1.0 : 1.0);
      if (bc[3] == BC_OUTFLOW) hv[c] = h[c] * outVel(s, h, src, src - NX);
    }
#undef PUT
  /* water-level boundary: the ghost surface held at the stage when it stands above the ghost's bed, with
   * the edge cell's velocity (solver-cpu.js:197-202) */
  for (int m = 0; m < s->nStage; m++) {
    int c = s->stageG[m], e = s->stageS[m];
    double hs = s->stageLevel - s->z[c];
    if (!(hs > s->p.hdry)) continue;
    double u = h[e] > s->p.hdry ? hu[e] / h[e] : 0, v = h[e] > s->p.hdry ? hv[e] / h[e] : 0;
    h[c] = hs; hu[c] = hs * u; hv[c] = hs * v;
  }
}

/* stageCells (solver-cpu.js:109-117): interior edge cells whose outward ghosts (both layers) see the stage */
void or_set_stage_cells(or_solver *s, const int *cells, int n) {
  free(s->stageG); free(s->stageS);
  s->stageG = (int *)calloc((size_t)(n ? 4 * GH * n : 1), sizeof(int));
  s->stageS = (int *)calloc((size_t)(n ? 4 * GH * n : 1), sizeof(int));
  s->nStage = 0;
  const int nx = s->p.nx, ny = s->p.ny, NX = s->NX;
  for (int q = 0; q < n; q++) {
    int i = cells[q] % nx, j = cells[q] / nx, e = (j + GH) * NX + i + GH;
#define ADD(ii, jj) do { s->stageG[s->nStage] = ((jj) + GH) * NX + (ii) + GH; s->stageS[s->nStage] = e; s->nStage++; } while (0)
    for (int g2 = 1; g2 <= GH; g2++) {
      if (i == 0) ADD(-g2, j);
      if (i == nx - 1) ADD(nx - 1 + g2, j);
      if (j == 0) ADD(i, -g2);
      if (j == ny - 1) ADD(i, ny - 1 + g2);
    }
#undef ADD
  }
}
void or_set_stage(or_solver *s, double z) { s->stageLevel = z; }

/* crest levels per barrier face (absolute; non-finite = not deployed), weir coefficient kept
 * (setCrestLevels, solver-cpu.js:136-138) */
void or_set_crest_levels(or_solver *s, const double *z) {
  for (int m = 0; m < s->nBar; m++) {
    double zc = isfinite(z[m]) ? z[m] : -INFINITY;
    if (s->barAxis[m]) s->CY[s->barC[m]] = zc; else s->CX[s->barC[m]] = zc;
  }
}

/* o.crest (solver-cpu.js:98-105): faces {i, j, axis, z, C, ground}; axis 0 = the face between (i-1, j) and
 * (i, j), axis 1 = between (i, j-1) and (i, j); books start empty (level -inf, tOver NaN) */
void or_set_crests(or_solver *s, const int *i, const int *j, const int *axis, const double *z, const double *C,
                   const double *ground, int n) {
  free(s->barC); free(s->barAxis); free(s->barGround); free(s->bar);
  s->nBar = n;
  s->barC = (int *)calloc((size_t)(n ? n : 1), sizeof(int));
  s->barAxis = (int *)calloc((size_t)(n ? n : 1), sizeof(int));
  s->barGround = zalloc((size_t)n);
  s->bar = zalloc((size_t)n * 9);
  for (int m = 0; m < n; m++) {
    s->barC[m] = (j[m] + GH) * s->NX + i[m] + GH;
    s->barAxis[m] = axis[m];
    s->barGround[m] = ground[m];
    s->bar[9 * m + 3] = -INFINITY;
    s->bar[9 * m + 8] = NAN;
    if (axis[m]) s->WY[s->barC[m]] = C[m]; else s->WX[s->barC[m]] = C[m];
  }
  or_set_crest_levels(s, z);
}
void or_read_barriers(const or_solver *s, double *out) { memcpy(out, s->bar, sizeof(double) * 9 * (size_t)s->nBar); }

/* o.ga (solver-cpu.js:96): Green-Ampt soil where perv > 0 (Fmax capped at 1e6); o.gi (:95): GI stores where
 * cap > 0; per interior cell */
void or_set_ga(or_solver *s, const double *psiD, const double *fmax, const double *perv) {
  for (int j = 0; j < s->p.ny; j++)
    for (int i = 0; i < s->p.nx; i++) {
      int c = (j + GH) * s->NX + i + GH, k = j * s->p.nx + i;
      if (perv[k] > 0) { s->gaPsiD[c] = psiD[k]; s->gaFmax[c] = dmin(fmax[k], 1e6); s->gaPerv[c] = perv[k]; }
    }
}
void or_set_gi(or_solver *s, const double *rate, const double *cap, const double *drain) {
  for (int j = 0; j < s->p.ny; j++)
    for (int i = 0; i < s->p.nx; i++) {
      int c = (j + GH) * s->NX + i + GH, k = j * s->p.nx + i;
      if (cap[k] > 0) { s->giRate[c] = rate[k]; s->giCap[c] = cap[k]; s->giDrain[c] = drain[k]; }
    }
}

static void prims(or_solver *s, const double *h, const double *hu, const double *hv) {
  const int N = s->N;
  const double hdry = s->p.hdry;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < N; c++) {
    double d = h[c];
    s->Hp[c] = d;
    s->Ep[c] = d + s->z[c];
    if (d > hdry) { s->Up[c] = hu[c] / d; s->Vp[c] = hv[c] / d; }
    else { s->Up[c] = 0; s->Vp[c] = 0; }
  }
}

/* MUSCL toward side sg along stride st; out = h, eta, un, ut (solver-cpu.js:216-233) */
static void recon(const or_solver *s, int c, int st, double sg, int isX, double *out) {
  const double *Hp = s->Hp, *Ep = s->Ep, *un = isX ? s->Up : s->Vp, *ut = isX ? s->Vp : s->Up;
  const double hfront = s->p.hfront;
  int fo = !s->p.secondOrder || Hp[c] < hfront;
  int a = c - st, b = c + st;
  int Ma = s->M[a], Mb = s->M[b];
  if (!fo && ((Ma && Hp[a] < hfront) || (Mb && Hp[b] < hfront))) fo = 1;
  /* F26: a film thinner than the bed's change across the cell is a cascade, reconstructed first order
   * (solver-cpu.js:222-224) */
  if (!fo && s->p.stepFlux && !s->p.lts && s->n[c] > 0 &&
      Hp[c] < 0.5 * fabs((Mb ? s->z[b] : s->z[c]) - (Ma ? s->z[a] : s->z[c])))
    fo = 1;
  if (fo) { out[0] = Hp[c]; out[1] = Ep[c]; out[2] = un[c]; out[3] = ut[c]; return; }
  double half = 0.5 * sg;
#define SL(q) lim(s, (q)[c] - (Ma ? (q)[a] : (q)[c]), (Mb ? (q)[b] : (q)[c]) - (q)[c])
  out[0] = Hp[c] + half * SL(Hp);
  out[1] = Ep[c] + half * SL(Ep);
  out[2] = un[c] + half * SL(un);
  out[3] = ut[c] + half * SL(ut);
#undef SL
}

/* HLLC in the face-normal frame (solver-cpu.js:579-602) */
static void hllc(double hL, double uL, double vL, double hR, double uR, double vR, double g, double hdry,
                 double *F) {
  int wetL = hL > hdry, wetR = hR > hdry;
  if (!wetL && !wetR) { F[0] = 0; F[1] = 0; F[2] = 0; return; }
  if (!wetL) { uL = 0; vL = 0; }
  if (!wetR) { uR = 0; vR = 0; }
  double cL = sqrt(g * hL), cR = sqrt(g * hR);
  double SL, SR;
  if (!wetL) { SL = uR - 2 * cR; SR = uR + cR; }
  else if (!wetR) { SL = uL - cL; SR = uL + 2 * cL; }
  else {
    double us = 0.5 * (uL + uR) + cL - cR;
    double cs = 0.5 * (cL + cR) + 0.25 * (uL - uR);
    SL = dmin(uL - cL, us - cs);
    SR = dmax(uR + cR, us + cs);
  }
  double FL0 = hL * uL, FL1 = hL * uL * uL + 0.5 * g * hL * hL;
  double FR0 = hR * uR, FR1 = hR * uR * uR + 0.5 * g * hR * hR;
  if (SL >= 0) { F[0] = FL0; F[1] = FL1; F[2] = FL0 * vL; return; }
  if (SR <= 0) { F[0] = FR0; F[1] = FR1; F[2] = FR0 * vR; return; }
  double den = SR - SL;
  double H0 = (SR * FL0 - SL * FR0 + SL * SR * (hR - hL)) / den;
  double H1 = (SR * FL1 - SL * FR1 + SL * SR * (hR * uR - hL * uL)) / den;
  double dSM = hR * (uR - SR) - hL * (uL - SL);
  double SM = fabs(dSM) > 1e-300 ? (SL * hR * (uR - SR) - SR * hL * (uL - SL)) / dSM : 0;
  F[0] = H0; F[1] = H1; F[2] = H0 * (SM >= 0 ? vL : vR);
}

/* flow over a barrier crest zc: broad-crested weir q = C H1^1.5 with Villemonte submergence
 * (1 - (H2/H1)^1.5)^0.385, capped so one step cannot more than half-close the level difference; the
 * momentum carries the critical velocity over the crest, the pressure is the mean of the two sides'
 * depths over the crest (_weir, solver-cpu.js:277-287) */
static void weir(const or_solver *s, int a, int b, int isX, double zc, double cw, double hLs, double hRs, double *F) {
  const double *Hp = s->Hp, *Ep = s->Ep, g = s->p.g, hdry = s->p.hdry, dx = s->p.dx;
  int left = Ep[a] >= Ep[b], up = left ? a : b, dn = left ? b : a;
  double pres = 0.25 * g * (hLs * hLs + hRs * hRs), H1 = Ep[up] - zc;
  F[0] = 0; F[1] = pres; F[2] = 0;
  if (H1 <= hdry || Hp[up] <= hdry) return;
  double r = dmax(0, Ep[dn] - zc) / H1;
  if (r >= 1) return;
  double q = cw * pow(H1, 1.5) * pow(1 - pow(r, 1.5), 0.385);
  if (s->dtCur > 0) q = dmin(q, (Ep[up] - Ep[dn]) * dx / (4 * s->dtCur));
  double un = q / dmax(2.0 / 3.0 * H1, hdry), F0 = left ? q : -q;
  F[0] = F0; F[1] = q * un + pres; F[2] = F0 * (isX ? s->Vp[up] : s->Up[up]);
}

/* face between a (left/below) and b with crest zc (-INFINITY: none) and weir coefficient cw;
 * F = [F0, F1, F2, corrL, corrR, hL, zL, hR, zR] (solver-cpu.js:237-267) */
static void face(const or_solver *s, int a, int b, int st, int isX, double *F, double zc, double cw) {
  int Ma = s->M[a], Mb = s->M[b];
  if (!Ma && !Mb) { for (int k = 0; k < 9; k++) F[k] = 0; return; }
  double r[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (Ma) recon(s, a, st, +1, isX, r);
  if (Mb) recon(s, b, st, -1, isX, r + 4);
  if (!Mb) { r[4] = r[0]; r[5] = r[1]; r[6] = -r[2]; r[7] = r[3]; }
  if (!Ma) { r[0] = r[4]; r[1] = r[5]; r[2] = -r[6]; r[3] = r[7]; }
  double hL = r[0], zL = r[1] - r[0], uL = r[2], vL = r[3];
  double hR = r[4], zR = r[5] - r[4], uR = r[6], vR = r[7];
  double zs = dmax(dmax(zL, zR), zc);
  double hLs = dmax(0, r[1] - zs), hRs = dmax(0, r[5] - zs);
  const double g = s->p.g, hdry = s->p.hdry;
  int isWeir = Ma && Mb && zc > dmax(s->z[a], s->z[b]);
  /* F26 step face (solver-cpu.js:250-254): a wet film on one side whose reconstructed depth over the
   * other side's higher bed is zero passes its own discharge, upwind (_stepFlux, :269-273) */
  int cas = 0;
  if (s->p.stepFlux && !s->p.lts && !isWeir && Ma && Mb && s->n[a] > 0 && s->n[b] > 0) {
    if (hLs > hdry && hRs <= 0 && hR > hdry) cas = 1;
    else if (hRs > hdry && hLs <= 0 && hL > hdry) cas = -1;
  }
  if (isWeir) {
    weir(s, a, b, isX, zc, cw, hLs, hRs, F);
  } else if (cas == 1) {
    double u = dmax(uL, 0), q = hLs * u;
    F[0] = q; F[1] = q * u + 0.5 * g * hLs * hLs; F[2] = q * vL;
  } else if (cas == -1) {
    double u = dmin(uR, 0), q = hRs * u;
    F[0] = q; F[1] = q * u + 0.5 * g * hRs * hRs; F[2] = q * vR;
  } else {
    hllc(hLs, uL, vL, hRs, uR, vR, g, hdry, F);
  }
  F[3] = 0.5 * g * (hL * hL - hLs * hLs);
  F[4] = 0.5 * g * (hR * hR - hRs * hRs);
  /* the lower cell feels the whole step (solver-cpu.js:261-264) */
  if (cas == 1) F[4] = 0.5 * g * hR * hR + g * hR * (zs - r[5]);
  else if (cas == -1) F[3] = 0.5 * g * hL * hL + g * hL * (zs - r[1]);
  F[5] = hL; F[6] = zL; F[7] = hR; F[8] = zR;
  if (!Mb) F[4] = 0;
  if (!Ma) F[3] = 0;
}

/* spatial operator L(U) into dh/dhu/dhv; returns the boundary outflow rate (m^3/s) (solver-cpu.js:291-381) */
static double opL(or_solver *s, double *h, double *hu, double *hv, double dt) {
  const int NX = s->NX, NY = s->NY, lev = s->lev;
  double *FX = s->FX, *FY = s->FY, *TH = s->TH;
  const double dx = s->p.dx, g = s->p.g;
  s->dtCur = dt;
  fillGhost(s, h, hu, hv);
  prims(s, h, hu, hv);
#define ON(a, b) (lev < 0 || (lvc(s, a) < lvc(s, b) ? lvc(s, a) : lvc(s, b)) == lev)
#define WF(a, b) (ON(a, b) ? 1.0 : 0.0)
#define NEAR(c) (lev < 0 || lvc(s, c) == lev || lvc(s, c) == lev + 1)
#pragma omp parallel for schedule(static)
  for (int j = GH; j < NY - GH; j++)
    for (int i = GH; i <= NX - GH; i++) {
      int b = j * NX + i;
      if (ON(b - 1, b)) face(s, b - 1, b, 1, 1, FX + (size_t)b * 9, s->CX[b], s->WX[b]);
    }
#pragma omp parallel for schedule(static)
  for (int j = GH; j <= NY - GH; j++)
    for (int i = GH; i < NX - GH; i++) {
      int b = j * NX + i;
      if (ON(b - NX, b)) face(s, b - NX, b, NX, 0, FY + (size_t)b * 9, s->CY[b], s->WY[b]);
    }
  /* draining-time limiter */
  const double area = dx * dx;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < s->N; c++) TH[c] = 1;
#pragma omp parallel for schedule(static)
  for (int j = GH; j < NY - GH; j++)
    for (int i = GH; i < NX - GH; i++) {
      int c = j * NX + i;
      if (!s->M[c] || !NEAR(c)) continue;
      double fw = FX[(size_t)c * 9] * WF(c - 1, c), fe = FX[(size_t)(c + 1) * 9] * WF(c, c + 1);
      double fs = FY[(size_t)c * 9] * WF(c - NX, c), fn = FY[(size_t)(c + NX) * 9] * WF(c, c + NX);
      double out = ((fe > 0 ? fe : 0) - (fw < 0 ? fw : 0) + (fn > 0 ? fn : 0) - (fs < 0 ? fs : 0)) * dx * dt;
      double vol = (lev >= 0 && (lvc(s, c) == lev + 1 || !s->stage)) ? dmax(0, h[c] + s->aH[c]) * area : h[c] * area;
      if (out > vol) TH[c] = vol / out;
    }
#pragma omp parallel for schedule(static)
  for (int j = GH; j < NY - GH; j++)
    for (int i = GH; i <= NX - GH; i++) {
      int b = j * NX + i;
      double *f = FX + (size_t)b * 9;
      double t = TH[f[0] > 0 ? b - 1 : b];
      if (t < 1 && ON(b - 1, b)) { f[0] *= t; f[1] *= t; f[2] *= t; }
    }
#pragma omp parallel for schedule(static)
  for (int j = GH; j <= NY - GH; j++)
    for (int i = GH; i < NX - GH; i++) {
      int b = j * NX + i;
      double *f = FY + (size_t)b * 9;
      double t = TH[f[0] > 0 ? b - NX : b];
      if (t < 1 && ON(b - NX, b)) { f[0] *= t; f[1] *= t; f[2] *= t; }
    }
  /* barrier books (solver-cpu.js:327-343): the theta-scaled flux each barrier face passes this stage, booked
   * while the barrier is up and water stands 2 cm over its crest; in the first stage the peak rate, the
   * highest water and the load, each with the step's start time */
  {
    const double wgt = s->p.secondOrder ? 0.5 : 1;
    for (int m = 0; m < s->nBar; m++) {
      int b = s->barC[m], a = b - (s->barAxis[m] ? NX : 1);
      double crest = (s->barAxis[m] ? s->CY : s->CX)[b], *B = s->bar + 9 * m;
      if ((s->Hp[a] <= 0 && s->Hp[b] <= 0) || crest == -INFINITY || !ON(a, b)) continue;
      double ea = s->Hp[a] > 0.005 ? s->Ep[a] : -INFINITY, eb = s->Hp[b] > 0.005 ? s->Ep[b] : -INFINITY;
      double w = dmax(ea, eb);
      double q = (s->barAxis[m] ? FY : FX)[(size_t)b * 9];
      int over = w > crest + 0.02;
      if (over) {
        double v = q * dx * dt * wgt;
        if (v > 0) B[0] += v; else B[1] -= v;
        if (isnan(B[8]) && v != 0) B[8] = s->t;
      }
      if (s->stage) continue;
      if (over && fabs(q) > B[2]) { B[2] = fabs(q); B[5] = s->t; }
      if (w > B[3]) { B[3] = w; B[6] = s->t; }
      double gr = s->barGround[m];
      double dA = dmax(0, dmin(ea, crest) - gr), dB = dmax(0, dmin(eb, crest) - gr);
      double load = isfinite(gr) ? 0.5 * g * fabs(dA * dA - dB * dB) : 0;
      if (load > B[4]) { B[4] = load; B[7] = s->t; }
    }
  }
  const double inv = 1 / dx;
#pragma omp parallel for schedule(static)
  for (int j = GH; j < NY - GH; j++)
    for (int i = GH; i < NX - GH; i++) {
      int c = j * NX + i;
      if (!s->M[c] || !NEAR(c)) { s->dh[c] = s->dhu[c] = s->dhv[c] = 0; continue; }
      const double *w = FX + (size_t)c * 9, *e = FX + (size_t)(c + 1) * 9;
      const double *sth = FY + (size_t)c * 9, *nth = FY + (size_t)(c + NX) * 9;
      double hW = w[7], zW = w[8], hE = e[5], zE = e[6];
      double hS = sth[7], zS = sth[8], hN = nth[5], zN = nth[6];
      if (lev < 0) {
        s->dh[c] = -inv * (e[0] - w[0] + nth[0] - sth[0]);
        s->dhu[c] = -inv * ((e[1] + e[3]) - (w[1] + w[4]) + 0.5 * g * (hW + hE) * (zE - zW) + nth[2] - sth[2]);
        s->dhv[c] = -inv * ((nth[1] + nth[3]) - (sth[1] + sth[4]) + 0.5 * g * (hS + hN) * (zN - zS) + e[2] - w[2]);
        continue;
      }
      double kw = WF(c - 1, c), ke = WF(c, c + 1), ks = WF(c - NX, c), kn = WF(c, c + NX);
      double own = lvc(s, c) == lev ? 1 : 0;
      s->dh[c] = -inv * (ke * e[0] - kw * w[0] + kn * nth[0] - ks * sth[0]);
      s->dhu[c] = -inv * (ke * (e[1] + e[3]) - kw * (w[1] + w[4]) + own * 0.5 * g * (hW + hE) * (zE - zW) +
                          kn * nth[2] - ks * sth[2]);
      s->dhv[c] = -inv * (kn * (nth[1] + nth[3]) - ks * (sth[1] + sth[4]) + own * 0.5 * g * (hS + hN) * (zN - zS) +
                          ke * e[2] - kw * w[2]);
    }
  /* boundary mass fluxes; with LTS an edge face steps with its edge cell */
#define EK(c) ((lev < 0 || lvc(s, c) == lev) ? 1.0 : 0.0)
  double qout = 0;
  for (int j = GH; j < NY - GH; j++) {
    qout -= EK(j * NX + GH) * FX[(size_t)(j * NX + GH) * 9] * dx;
    qout += EK(j * NX + NX - GH - 1) * FX[(size_t)(j * NX + NX - GH) * 9] * dx;
  }
  for (int i = GH; i < NX - GH; i++) {
    qout -= EK(GH * NX + i) * FY[(size_t)(GH * NX + i) * 9] * dx;
    qout += EK((NY - GH - 1) * NX + i) * FY[(size_t)((NY - GH) * NX + i) * 9] * dx;
  }
#undef EK
#undef ON
#undef WF
#undef NEAR
  return qout;
}

/* CFL step over wet, non-wall interior cells; +inf when all dry (solver-cpu.js:383-394) */
double or_compute_dt(or_solver *s) {
  const int NX = s->NX, NY = s->NY;
  const double g = s->p.g, hdry = s->p.hdry;
  double smax = 0;
#pragma omp parallel for schedule(static) reduction(max : smax)
  for (int j = GH; j < NY - GH; j++)
    for (int i = GH; i < NX - GH; i++) {
      int c = j * NX + i;
      double d = s->h[c];
      if (!s->M[c] || d <= hdry) continue;
      double cel = sqrt(g * d);
      double sp = fabs(s->hu[c] / d) + fabs(s->hv[c] / d) + 2 * cel;
      if (sp > smax) smax = sp;
    }
  return smax > 0 ? s->p.cfl * s->p.dx / smax : INFINITY;
}

/* clip negative depths (booked), zero momentum where dry (solver-cpu.js:396-404); sel: -1 all, else level */
static void fixcells(or_solver *s, int sel) {
  const int NX = s->NX, NY = s->NY;
  const double hdry = s->p.hdry;
#pragma omp parallel for schedule(static)
  for (int j = GH; j < NY - GH; j++) {
    double cl = 0;
    for (int i = GH; i < NX - GH; i++) {
      int c = j * NX + i;
      if (!s->M[c]) continue;
      if (sel >= 0 && lvc(s, c) != sel) continue;
      if (s->h[c] < 0) { cl -= s->h[c]; s->h[c] = 0; }
      if (s->h[c] <= hdry) { s->hu[c] = 0; s->hv[c] = 0; }
    }
    s->rowA[j] = cl;
  }
  double clipped = 0;
  for (int j = GH; j < NY - GH; j++) clipped += s->rowA[j];
  s->volClipped += clipped * s->p.dx * s->p.dx;
}

/* semi-implicit Manning friction, rain, infiltration, max depth (solver-cpu.js:456-483); sel as fixcells */
/* hazard standards (hazard.js:30-66): ARR 2019 / AIDR 7-3 H1-H6 and FD2320 HR = d(v + 0.5) + DF */
#define HZ_WET 0.01
#define HZ_VCUT 0.05
#define HZ_THR 0.15
int or_hz_arr(double d, double v) {
  double dv = d * v;
  if (dv <= 0.3 && d <= 0.3 && v <= 2) return 1;
  if (dv <= 0.6 && d <= 0.5 && v <= 2) return 2;
  if (dv <= 0.6 && d <= 1.2 && v <= 2) return 3;
  if (dv <= 1.0 && d <= 2.0 && v <= 2) return 4;
  if (dv <= 4.0 && d <= 4.0 && v <= 4) return 5;
  return 6;
}
int or_hz_hr(double hr) { return hr < 0.75 ? 1 : hr < 1.25 ? 2 : hr < 2 ? 3 : 4; }
/* momentum flux per metre of width, kN/m: rho (d v^2 + g d^2 / 2) / 1000 with rho = 1000 kg/m^3
 * (hazard.js:71) */
double or_hz_force(double d, double v, double g) { return d * v * v + 0.5 * g * d * d; }

/* HazardTracker.cell (hazard.js:81-96) for interior cell k after a step of length dt ending at t */
static void hazard_cell(or_solver *s, int k, double d, double v, double t, double dt) {
  if (!(d >= HZ_WET)) return;
  double dv = d * v;
  if (dv > s->hzDv[k]) s->hzDv[k] = dv;
  double f = or_hz_force(d, v, s->p.g);
  if (f > s->hzForce[k]) s->hzForce[k] = f;
  if (d > HZ_THR) { s->hzDur[k] += dt; if (isnan(s->hzArr[k])) s->hzArr[k] = t; }
  if (d > HZ_VCUT) { if (v > s->hzVmax[k]) s->hzVmax[k] = v; if (isnan(s->hzWet[k])) s->hzWet[k] = t; }
  int c[3] = {or_hz_arr(d, v), or_hz_hr(d * (v + 0.5)), or_hz_hr(d * (v + 0.5) + (d > 0.25 ? 1 : 0))};
  for (int q = 0; q < 3; q++)
    if (c[q] > s->hzCls[3 * k + q]) s->hzCls[3 * k + q] = (uint8_t)c[q];
}

/* the tracker on one cell, for known-answer tests of its rules */
void or_hazard_cell(or_solver *s, int k, double d, double v, double t, double dt) { hazard_cell(s, k, d, v, t, dt); }

static void sources(or_solver *s, double dt, int sel, double tEnd) {
  const int NX = s->NX, NY = s->NY;
  const double g = s->p.g, hdry = s->p.hdry, r0 = s->rainRate, A = s->p.dx * s->p.dx;
#pragma omp parallel for schedule(static)
  for (int j = GH; j < NY - GH; j++) {
    double rv = 0, iv = 0;
    for (int i = GH; i < NX - GH; i++) {
      int c = j * NX + i;
      if (!s->M[c]) continue;
      if (sel >= 0 && lvc(s, c) != sel) continue;
      double d = s->h[c];
      if (d > hdry && s->n[c] > 0) {
        double u = s->hu[c] / d, v = s->hv[c] / d;
        double a = dt * g * s->n[c] * s->n[c] / pow(d, 4.0 / 3.0), sp = sqrt(u * u + v * v);
        /* F26: fully implicit in the new speed (solver-cpu.js:465-467) */
        double k = s->p.implicitFriction ? 0.5 * (1 + sqrt(1 + 4 * a * sp)) : 1 + a * sp;
        s->hu[c] /= k; s->hv[c] /= k;
      }
      double r = r0 * (s->rainField ? s->rainField[c] : 1) * dt;
      double cap = s->infil[c] * dt;
      /* Green-Ampt: capacity K (1 + psi dtheta / F), none once the soil above the water table is full
       * (solver-cpu.js:466) */
      if (s->gaPerv[c] > 0)
        cap = s->gaF[c] >= s->gaFmax[c] ? 0
                                        : dmin(cap * (1 + s->gaPsiD[c] / dmax(s->gaF[c], 1e-3)),
                                               (s->gaFmax[c] - s->gaF[c]) * s->gaPerv[c]);
      double inf = dmin(cap, d + r);
      if (s->gaPerv[c] > 0) s->gaF[c] += inf / s->gaPerv[c];
      /* GI store: drains linearly, then takes what it can at its rate up to its capacity (:469-472) */
      if (s->giCap[c] > 0) {
        double st = dmax(0, s->giStore[c] - s->giDrain[c] * dt);
        double g2 = dmin(dmin(s->giRate[c] * dt, d + r - inf), s->giCap[c] - st);
        s->giStore[c] = st + dmax(0, g2);
        if (g2 > 0) inf += g2;
      }
      s->h[c] = dmax(0, d + r - inf);
      rv += r;
      iv += d + r - s->h[c];
      if (s->h[c] <= hdry) { s->hu[c] = 0; s->hv[c] = 0; }
      if (s->h[c] > s->maxDepth[c]) s->maxDepth[c] = s->h[c];
    }
    s->rowA[j] = rv; s->rowB[j] = iv;
  }
  double rv = 0, iv = 0;
  for (int j = GH; j < NY - GH; j++) { rv += s->rowA[j]; iv += s->rowB[j]; }
  s->volRain += rv * A;
  s->volInfil += iv * A;
  /* hazard products after the sources, the step ending at tEnd (solver-cpu.js:484-487) */
#pragma omp parallel for schedule(static)
  for (int j = GH; j < NY - GH; j++)
    for (int i = GH; i < NX - GH; i++) {
      int c = j * NX + i;
      if (!s->M[c]) continue;
      if (sel >= 0 && lvc(s, c) != sel) continue;
      double d = s->h[c];
      double v = d >= HZ_VCUT ? sqrt(s->hu[c] * s->hu[c] + s->hv[c] * s->hv[c]) / d : 0;
      hazard_cell(s, (j - GH) * s->p.nx + (i - GH), d, v, tEnd, dt);
    }
}

/* inlet exchange over dt (exec-cpu.mjs:46-55); sel: -1 all inlets, else inlets whose cell is at that level */
static void exchange(or_solver *s, double dt, int sel) {
  const double A = s->p.dx * s->p.dx;
  for (int m = 0; m < s->nIn; m++) {
    int c = s->inC[m];
    if (sel >= 0 && lvc(s, c) != sel) continue;
    double dd = dmin(s->inQ[m] * dt / A, dmax(s->h[c], 0));
    s->h[c] -= dd;
    if (s->h[c] <= 0) s->h[c] = 0;
    if (s->h[c] <= s->p.hdry) { s->hu[c] = 0; s->hv[c] = 0; }
    s->inV[m] += dd * A;
    s->exchanged += dd * A;
  }
}

#define INTERIOR_M(body)                                   \
  _Pragma("omp parallel for schedule(static)")             \
  for (int j = GH; j < s->NY - GH; j++)                    \
    for (int i = GH; i < s->NX - GH; i++) {                \
      int c = j * s->NX + i;                               \
      if (!s->M[c]) continue;                              \
      body                                                 \
    }

static double cycle(or_solver *s, double dtMax);

/* friction alone over dt, implicit form, on the interior open cells (sel -1) or those of level sel
 * (_friction, solver-cpu.js:443-452) */
static void friction(or_solver *s, double dt, int sel) {
  const double g = s->p.g, hdry = s->p.hdry;
  double *h = s->h, *hu = s->hu, *hv = s->hv;
#pragma omp parallel for schedule(static)
  for (int j = GH; j < s->NY - GH; j++)
    for (int i = GH; i < s->NX - GH; i++) {
      int c = j * s->NX + i;
      if (!s->M[c] || (sel >= 0 && lvc(s, c) != sel)) continue;
      double d = h[c];
      if (d <= hdry || !(s->n[c] > 0)) continue;
      double a = dt * g * s->n[c] * s->n[c] / pow(d, 4.0 / 3.0);
      double sp = hypot(hu[c], hv[c]) / d;
      double k = 0.5 * (1 + sqrt(1 + 4 * a * sp));
      hu[c] /= k;
      hv[c] /= k;
    }
}
/* the second stage's fluxes see the predictor after friction; its update keeps the predictor
 * (solver-cpu.js:419-423, 542-545) */
static void keep_momentum(or_solver *s, int save) {
  if (save) {
    memcpy(s->keepU, s->hu, sizeof(double) * s->N);
    memcpy(s->keepV, s->hv, sizeof(double) * s->N);
  } else {
    memcpy(s->hu, s->keepU, sizeof(double) * s->N);
    memcpy(s->hv, s->keepV, sizeof(double) * s->N);
  }
}

/* one global step (solver-cpu.js:406-441) */
double or_step(or_solver *s, double dtMax) {
  if (s->L) return cycle(s, dtMax);
  double dtc = or_compute_dt(s);
  double dt = dmin(isfinite(dtc) ? dtc : s->p.dtDry, dtMax);
  double *h = s->h, *hu = s->hu, *hv = s->hv;
  if (s->p.secondOrder) {
    memcpy(s->h0, h, sizeof(double) * s->N);
    memcpy(s->hu0, hu, sizeof(double) * s->N);
    memcpy(s->hv0, hv, sizeof(double) * s->N);
    s->stage = 0;
    double q1 = opL(s, h, hu, hv, dt);
    INTERIOR_M(h[c] += dt * s->dh[c]; hu[c] += dt * s->dhu[c]; hv[c] += dt * s->dhv[c];)
    fixcells(s, -1);
    if (s->p.implicitFriction) { keep_momentum(s, 1); friction(s, dt, -1); }
    s->stage = 1;
    double q2 = opL(s, h, hu, hv, dt);
    if (s->p.implicitFriction) keep_momentum(s, 0);
    INTERIOR_M(h[c] = 0.5 * (s->h0[c] + h[c] + dt * s->dh[c]); hu[c] = 0.5 * (s->hu0[c] + hu[c] + dt * s->dhu[c]);
               hv[c] = 0.5 * (s->hv0[c] + hv[c] + dt * s->dhv[c]);)
    fixcells(s, -1);
    s->volOutflow += dt * 0.5 * (q1 + q2);
  } else {
    s->stage = 0;
    double q = opL(s, h, hu, hv, dt);
    INTERIOR_M(h[c] += dt * s->dh[c]; hu[c] += dt * s->dhu[c]; hv[c] += dt * s->dhv[c];)
    fixcells(s, -1);
    s->volOutflow += dt * q;
  }
  sources(s, dt, -1, s->t + dt);
  if (s->nIn) exchange(s, dt, -1);
  s->t += dt;
  s->steps++;
  return dt;
}

/* one level-m step of an LTS cycle, dt = 2^m * base step (solver-cpu.js:529-558) */
static void record(or_solver *s, int m, double dt, double tEnd) {
  double *h = s->h, *hu = s->hu, *hv = s->hv;
  const double hdry = s->p.hdry;
  (void)hdry;
  long ncells = 0;
  for (int j = GH; j < s->NY - GH; j++)
    for (int i = GH; i < s->NX - GH; i++) {
      int c = j * s->NX + i;
      if (s->M[c] && lvc(s, c) == m) ncells++;
    }
  const double w = s->p.secondOrder ? 0.5 * dt : dt;
#define OWN(c) (lvc(s, c) == m)
#define COARSE(c) (lvc(s, c) == m + 1)
#define GATHER() INTERIOR_M(if (COARSE(c)) { s->aH[c] += w * s->dh[c]; s->aHu[c] += w * s->dhu[c]; s->aHv[c] += w * s->dhv[c]; })
  s->lev = m;
  s->records++;
  s->cellSteps += ncells;
  if (s->p.secondOrder) {
    INTERIOR_M(if (OWN(c)) { s->h0[c] = h[c] + s->aH[c]; s->hu0[c] = hu[c] + s->aHu[c]; s->hv0[c] = hv[c] + s->aHv[c]; })
    s->stage = 0;
    double q1 = opL(s, h, hu, hv, dt);
    GATHER();
    INTERIOR_M(if (OWN(c)) { h[c] = s->h0[c] + dt * s->dh[c]; hu[c] = s->hu0[c] + dt * s->dhu[c]; hv[c] = s->hv0[c] + dt * s->dhv[c]; })
    fixcells(s, m);
    if (s->p.implicitFriction) { keep_momentum(s, 1); friction(s, dt, m); }
    s->stage = 1;
    double q2 = opL(s, h, hu, hv, dt);
    GATHER();
    if (s->p.implicitFriction) keep_momentum(s, 0);
    INTERIOR_M(if (OWN(c)) {
      h[c] = 0.5 * (s->h0[c] + h[c] + dt * s->dh[c]);
      hu[c] = 0.5 * (s->hu0[c] + hu[c] + dt * s->dhu[c]);
      hv[c] = 0.5 * (s->hv0[c] + hv[c] + dt * s->dhv[c]);
    })
    fixcells(s, m);
    s->volOutflow += dt * 0.5 * (q1 + q2);
  } else {
    s->stage = 0;
    double q = opL(s, h, hu, hv, dt);
    GATHER();
    INTERIOR_M(if (OWN(c)) { h[c] += s->aH[c] + dt * s->dh[c]; hu[c] += s->aHu[c] + dt * s->dhu[c]; hv[c] += s->aHv[c] + dt * s->dhv[c]; })
    fixcells(s, m);
    s->volOutflow += dt * q;
  }
  INTERIOR_M(if (OWN(c)) { s->aH[c] = 0; s->aHu[c] = 0; s->aHv[c] = 0; })
  s->lev = -1;
  sources(s, dt, m, tEnd);
  if (s->nIn) exchange(s, dt, m);
#undef OWN
#undef COARSE
#undef GATHER
}

/* one LTS cycle of 2^L base steps (solver-cpu.js:504-527) */
static double cycle(or_solver *s, double dtMax) {
  const int L = s->L, P = 1 << L, tx = s->tilesX, ty = s->tilesY, nt = tx * ty;
  const double g = s->p.g, hdry = s->p.hdry;
  double *sp = s->tsp;
  int *LV = s->LV;
  for (int t = 0; t < nt; t++) sp[t] = 0;
  double smax = 0;
  for (int j = GH; j < s->NY - GH; j++)
    for (int i = GH; i < s->NX - GH; i++) {
      int c = j * s->NX + i;
      double d = s->h[c];
      if (!s->M[c] || d <= hdry) continue;
      double v = fabs(s->hu[c] / d) + fabs(s->hv[c] / d) + 2 * sqrt(g * d);
      int t = s->TI[c];
      if (v > sp[t]) sp[t] = v;
      if (v > smax) smax = v;
    }
  double D = dmin(smax > 0 ? P * LTS_SAFE * s->p.cfl * s->p.dx / smax : s->p.dtDry, dtMax), dl = D / P;
  for (int t = 0; t < nt; t++) {
    int x = t % tx, y = t / tx;
    double v = 0;
    for (int b = (y - 1 > 0 ? y - 1 : 0); b <= (y + 1 < ty - 1 ? y + 1 : ty - 1); b++)
      for (int a = (x - 1 > 0 ? x - 1 : 0); a <= (x + 1 < tx - 1 ? x + 1 : tx - 1); a++) v = dmax(v, sp[b * tx + a]);
    if (v > 0) {
      double lf = floor(log2(LTS_SAFE * s->p.cfl * s->p.dx / v / dl));
      double mm = dmax(0, dmin((double)L, lf));
      LV[t] = (int)mm;
    } else LV[t] = L;
  }
  for (int p = 0; p < L; p++)
    for (int t = 0; t < nt; t++) {
      int x = t % tx, y = t / tx;
      if (x > 0 && LV[t - 1] + 1 < LV[t]) LV[t] = LV[t - 1] + 1;
      if (x < tx - 1 && LV[t + 1] + 1 < LV[t]) LV[t] = LV[t + 1] + 1;
      if (y > 0 && LV[t - tx] + 1 < LV[t]) LV[t] = LV[t - tx] + 1;
      if (y < ty - 1 && LV[t + tx] + 1 < LV[t]) LV[t] = LV[t + tx] + 1;
    }
  double t0 = s->t;
  for (int k = 0; k < P; k++)
    for (int m = 0; m <= L; m++)
      if ((k + 1) % (1 << m) == 0) record(s, m, dl * (1 << m), k == P - 1 ? t0 + D : t0 + (k + 1) * dl);
  s->t = t0 + D;
  s->steps++;
  return D;
}

/* run to tEnd (solver-cpu.js:560-562) */
void or_run(or_solver *s, double tEnd) {
  while (s->t < tEnd - 1e-12) or_step(s, tEnd - s->t);
}
/* at most k steps toward tEnd; returns the steps taken */
long or_run_steps(or_solver *s, double tEnd, long k) {
  long n = 0;
  while (n < k && s->t < tEnd - 1e-12) { or_step(s, tEnd - s->t); n++; }
  return n;
}

/* hazard field: 0 d*v, 1 force, 2 duration, 3 arrival (NaN never), 4 peak speed, 5 first time deeper
 * than VCUT (NaN never), 6 ARR class, 7 FD2320 class, 8 FD2320 urban class; interior nx*ny */
void or_get_hazard(const or_solver *s, int which, double *out) {
  size_t n = (size_t)s->p.nx * s->p.ny;
  const double *src[6] = {s->hzDv, s->hzForce, s->hzDur, s->hzArr, s->hzVmax, s->hzWet};
  for (size_t k = 0; k < n; k++) out[k] = which < 6 ? src[which][k] : (double)s->hzCls[3 * k + (which - 6)];
}

/* field: 0 h, 1 hu, 2 hv, 3 maxDepth, 4 z, 5 aH; interior nx*ny row-major */
void or_get_field(const or_solver *s, int which, double *out) {
  const double *src = which == 0 ? s->h : which == 1 ? s->hu : which == 2 ? s->hv : which == 3 ? s->maxDepth
                    : which == 4 ? s->z : which == 6 ? s->gaF : which == 7 ? s->giStore : s->aH;
  for (int j = 0; j < s->p.ny; j++)
    for (int i = 0; i < s->p.nx; i++) out[j * s->p.nx + i] = src ? src[(j + GH) * s->NX + i + GH] : 0;
}
/* LTS tile levels of the last cycle (tilesX*tilesY) */
int or_get_levels(const or_solver *s, int *out) {
  if (!s->L) return 0;
  for (int t = 0; t < s->tilesX * s->tilesY; t++) out[t] = s->LV[t];
  return s->tilesX * s->tilesY;
}

/* books: [t, steps, volume, initial, rain, infil, outflow, clipped, exchanged, expected, rel, cellSteps, records]
 * expected and rel as massBalance (solver-cpu.js:564-568) with the exchanged volume of CpuSim.readState */
void or_books(const or_solver *s, double *o) {
  double v = or_volume(s);
  double expected = s->volInitial + s->volRain - s->volInfil - s->volOutflow - s->exchanged;
  double scale = dmax(1e-30, dmax(fabs(expected), s->volInitial + s->volRain + dmax(0, -s->exchanged)));
  o[0] = s->t; o[1] = (double)s->steps; o[2] = v; o[3] = s->volInitial; o[4] = s->volRain; o[5] = s->volInfil;
  o[6] = s->volOutflow; o[7] = s->volClipped; o[8] = s->exchanged; o[9] = expected; o[10] = (v - expected) / scale;
  o[11] = (double)s->cellSteps; o[12] = (double)s->records;
}
/* book the initial volume again (a window restarted from a saved state) and set the clock */
void or_reset_books(or_solver *s, double t) {
  s->volInitial = or_volume(s);
  s->volRain = s->volInfil = s->volOutflow = s->volClipped = s->exchanged = 0;
  s->t = t;
  s->steps = 0;
  s->cellSteps = s->records = 0;
}

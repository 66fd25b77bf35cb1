// floodsim_hip runtime: device memory, work lists, the stepping loop and the C ABI of
// include/floodsim_hip.h. One translation unit with the kernels (sw_kernels.h).
//
// A pool owns any number of instances ("slots"), each a whole solver (the reference's GPUSolver,
// src/solver-gpu.js:13). The slots advanced together form a group: their descriptors, their
// wavefront strips (one work list) and their per-strip partials live in device arrays, so one
// launch of each kernel steps every instance of the group, whatever its size. A single solver
// (fs_handle) is a pool with one slot.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/floodsim_hip.h"
#include "sw_kernels.h"

using namespace fs;

namespace {

thread_local std::string g_create_error;

struct Slot {
  bool used = false;
  fs_params p{};
  int nx = 0, ny = 0, pitch = 0;
  double zref = 0.0;
  size_t bytes = 0;
  char *mem = nullptr;  // one device allocation for every per-cell array of the slot
  float *h = nullptr, *hu = nullptr, *hv = nullptr, *w = nullptr;
  float *h1 = nullptr, *hu1 = nullptr, *hv1 = nullptr, *w1 = nullptr;
  float *z = nullptr, *n = nullptr, *hmax = nullptr, *infil = nullptr, *rf = nullptr;
  uint8_t *mask = nullptr;
  Inlet *inl = nullptr;
  float *ispeed = nullptr;
  float4 *iapp = nullptr;
  int nin = 0;
  std::vector<int> inletCells;           // interior k = j*nx + i
  std::vector<uint8_t> hmask;            // host copy of the mask bits
  std::vector<float> hrf;                // host copy of the rain multiplier (empty = 1)
  int openCells = 0;
  int rows = 64;
  // dry-strip skipping (global stepping): per strip flags, rain and open-cell counts
  int sxn = 0, nstrips = 0;
  uint8_t *swet = nullptr, *sact = nullptr, *srain = nullptr;
  int32_t *sopen = nullptr;
  float4 *hzA = nullptr, *hzB = nullptr;   // hazard products (fs_params.hazard)
  // local time stepping
  int L = 0, tx = 0, ty = 0, nunits = 0;
  uint8_t *lv = nullptr, *lv0 = nullptr, *gflag = nullptr;
  uint64_t *urows = nullptr;
  double *rend = nullptr;
  unsigned *ticket = nullptr;
  int4 *plist = nullptr;
  float *tsp = nullptr, *tileRf = nullptr;
  unsigned *speedSlots = nullptr;
  int32_t *tileOpen = nullptr, *ucells = nullptr;
  float4 *st4 = nullptr;
  float4 *g4 = nullptr;
  float4 *bsx = nullptr, *bsy = nullptr;
  // flood barriers: one separate allocation (crest, bidx per cell; faces, ground, flux, books per barrier)
  char *bmem = nullptr;
  size_t bbytes = 0;
  int nbar = 0;
  float4 *crest = nullptr;
  int2 *bidx = nullptr;
  int4 *barFace = nullptr;
  float *barGround = nullptr, *barFlux = nullptr;
  double *barBook = nullptr;
  std::vector<float4> hcrest;            // host copy of the crest plane
  std::vector<int4> bars;                // per barrier: (i, j, axis, k of its right/upper cell)
  // Green-Ampt soil and GI stores: one separate allocation of up to two float4 planes
  char *smem = nullptr;
  size_t sbytes = 0;
  float4 *soil = nullptr, *gi = nullptr;
  // rain multiplier plane of a solver created without one and given one later (setRainField)
  char *rmem = nullptr;
  size_t rbytes = 0;
  int stageCells = 0;                    // cells flagged MASK_STAGE
  StepRec hrec{};                        // host mirror (refreshed after each launch group)
  double volInitial = 0.0;               // vol.initial (solver-gpu.js:309)
  double exchangedRead = 0.0;            // vol.exchanged: inlet volumes already read (solver-gpu.js:574)
  double dtSum = 0.0;                    // for dtAvg
  int64_t dtSteps = 0;
};

struct Group {
  std::vector<int> slots;
  DevInst *dinst = nullptr;
  int4 *dwork = nullptr;
  Part *p1 = nullptr, *p2 = nullptr;
  int nwork = 0, capWork = 0, capInst = 0;
  std::vector<DevInst> hinst;
  std::vector<int> withInlets;           // group indices of instances with inlets
  std::vector<int> withBarriers;         // group indices of instances with flood barriers
  bool anyFirstOrder = false, anySecondOrder = false;
  bool anyHazard = false;                // an instance tracks the hazard products
  bool anyStepFlux = false;              // an instance takes the step-flux rules (F26)
  unsigned kinds = 0;                    // second-stage kernel variants the group needs: bit 2*hazard + stepFlux
  bool lts = false;                      // a local-time-stepping instance (alone in its group)
  // captured K-step graphs, one per power-of-two K (index log2 K)
  hipGraphExec_t graphs[16] = {};
};

// What one step's (or cycle's) launch sequence needs: the device descriptors, work list and per-wavefront
// partials, and the host copies that pick the kernel variants. A Group gives one over its own arrays; the
// allocation-free advance builds one over the static scratch array and stack copies.
struct Launch {
  DevInst *dinst = nullptr;
  int4 *dwork = nullptr;
  Part *p1 = nullptr, *p2 = nullptr;
  int nwork = 0, ni = 0;
  const DevInst *hinst = nullptr;        // ni host copies
  const int *withInlets = nullptr;       // group indices of instances with inlets
  int nWithInlets = 0;
  const int *withBarriers = nullptr;     // group indices of instances with flood barriers
  int nWithBarriers = 0;
  unsigned kinds = 0;
  bool anyFirstOrder = false, anySecondOrder = false, lts = false;
  int slot0 = 0;                         // pool slot of the first instance (LTS timing histogram)
};

Launch view_of(const Group &G) {
  Launch L;
  L.dinst = G.dinst; L.dwork = G.dwork; L.p1 = G.p1; L.p2 = G.p2;
  L.nwork = G.nwork; L.ni = (int)G.slots.size();
  L.hinst = G.hinst.data();
  L.withInlets = G.withInlets.data(); L.nWithInlets = (int)G.withInlets.size();
  L.withBarriers = G.withBarriers.data(); L.nWithBarriers = (int)G.withBarriers.size();
  L.kinds = G.kinds; L.anyFirstOrder = G.anyFirstOrder; L.anySecondOrder = G.anySecondOrder; L.lts = G.lts;
  L.slot0 = G.slots.empty() ? 0 : G.slots[0];
  return L;
}

}  // namespace

struct fs_pool {
  size_t budget = 0, used = 0;
  int device = 0;                        // HIP device current at creation (every slot lives there)
  std::vector<Slot> slots;
  StepRec *drec = nullptr;               // device records, one per slot
  int capSlots = 0;
  hipStream_t stream = nullptr;
  Group grp;
  double *dtarget = nullptr;
  int capTarget = 0;
  bool timing = false, useGraphs = true;
  bool noSteps = false;                  // test switch (fs_pool_debug): launch groups run no steps
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  std::vector<hipEvent_t> tev;           // timing mode: 3 events per step (before, between, after the stages)
  int nTev = 0;                          // events recorded since the last fs_pool_kernel_ms
  double ms1 = 0.0, ms2 = 0.0;
  int64_t nTimed = 0;
  // timing mode, dominant kernel (second stage): launches, total ms, open cells it advanced
  struct Span { hipEvent_t a, b; int cycle, level; };
  std::vector<Span> spans;
  std::vector<hipEvent_t> spanEv;
  int nSpanEv = 0, nCycles = 0;
  int32_t *dhist = nullptr;              // per timed LTS cycle: own open cells per level
  int capHist = 0;
  double domMs = 0.0, domCells = 0.0;
  int64_t domLaunches = 0;
  float2 *pack = nullptr;                // frame readout staging (depth, speed per cell)
  size_t capPack = 0;
  std::string err;
};

struct fs_solver {
  fs_pool *pool = nullptr;
  int slot = -1;
};

namespace {

int fail(fs_pool *P, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (P) P->err = buf;
  else g_create_error = buf;
  return code;
}

#define HIPCHK(P, call)                                                                          \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess) return fail((P), FS_E_HIP, "%s: %s", #call, hipGetErrorString(e_));    \
  } while (0)

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// every plane is addressed with 32-bit byte offsets: 4-byte planes hold up to 2^30 cells, the 16-byte ones
// (LTS reconstructions, hazard products, Green-Ampt soil, GI stores, barrier crests) up to 2^28
constexpr int64_t kMax4BCells = (int64_t)1 << 30;
constexpr int64_t kMax16BCells = (int64_t)1 << 28;

// strip geometry: global stepping uses 58 owned columns (3 halo lanes each side); local time
// stepping uses 56 (4 halo lanes on the left) and units of rows in multiples of 8, at most 64, so that a
// unit holds whole 8x8 level tiles and its rows fit one 64-bit mask
int strip_own(const fs_params *p) { return p->lts ? 56 : kOwn; }
int strip_rows(const fs_params *p) {
  int r = p->rows > 0 ? std::max(8, p->rows) : 64;   // at least 8 rows (strip flag capacity)
  if (p->lts) r = std::min(64, (r + 7) / 8 * 8);
  return r;
}
int units_of(const fs_params *p) {
  return ((p->nx + strip_own(p) - 1) / strip_own(p)) * ((p->ny + strip_rows(p) - 1) / strip_rows(p));
}
// strip flag capacity of a global-stepping instance: room for the shortest strips (8 rows), so that
// the rows per strip can change after creation (fs_pool_set_rows)
constexpr int kMinRows = 8;
size_t strip_cap(const fs_params *p) {
  return (size_t)((p->nx + strip_own(p) - 1) / strip_own(p)) * (size_t)((p->ny + kMinRows - 1) / kMinRows);
}

size_t slot_bytes(const fs_params *p, bool infil, bool rf, int nin) {
  const size_t N = (size_t)p->nx * p->ny;
  size_t b = 0;
  if (p->lts) {
    const size_t nt = (size_t)((p->nx + 7) / 8) * ((p->ny + 7) / 8), nu = (size_t)units_of(p);
    b += 2 * align256(nt) + align256(nt * 4) * 3;           // lv, lv0, tsp, tileOpen, tileRf
    b += align256(kSpeedSlots * 4);                          // tile-speed block maxima
    b += 2 * align256(N * 16) + align256(N) + 2 * align256(N * 16);  // store, gathers + tst, flag, bsx, bsy
    b += align256(nu * 8 * (kMaxLevels + 1)) + align256(nu * 4 * (kMaxLevels + 1));   // urows, ucells
    b += align256(nu * kLtsUnitRuns * 16 * (kMaxLevels + 1));                         // plist
    b += align256(kLtsEndSlices * 3 * 8) + align256(4);                               // rend, ticket
  }
  if (!p->lts) b += 3 * align256(strip_cap(p)) + align256(strip_cap(p) * 4);  // strip flags
  if (p->hazard) b += 2 * align256(N * 16);  // hazard products
  b += 8 * align256(N * 4);               // A and B
  b += 3 * align256(N * 4);               // z, n, hmax
  b += align256(N);                       // mask
  if (infil) b += align256(N * 4);
  if (rf) b += align256(N * 4);
  b += align256((size_t)std::max(nin, 1) * sizeof(Inlet)) + align256((size_t)std::max(nin, 1) * 4);
  b += align256((size_t)std::max(nin, 1) * 16);
  return b;
}

int check_params(fs_pool *P, const fs_params *p) {
  if (!p) return fail(P, FS_E_ARG, "params is NULL");
  if (p->nx < 2 || p->ny < 2) return fail(P, FS_E_ARG, "nx and ny must be >= 2 (got %d x %d)", p->nx, p->ny);
  if ((int64_t)p->nx * p->ny > (p->lts || p->hazard ? kMax16BCells : kMax4BCells))
    return fail(P, FS_E_ARG, "grid too large: %d x %d (at most 2^30 cells, 2^28 with local time stepping or "
                "hazard products)", p->nx, p->ny);
  if (!(p->dx > 0)) return fail(P, FS_E_ARG, "dx must be > 0");
  if (!(p->cfl > 0) || !(p->hdry > 0) || !(p->dtDry > 0)) return fail(P, FS_E_ARG, "cfl, hdry, dtDry must be > 0");
  for (int k = 0; k < 4; ++k)
    if (p->bc[k] < 0 || p->bc[k] > 2) return fail(P, FS_E_ARG, "bc[%d] must be 0 wall, 1 open or 2 outflow", k);
  if (p->lts < 0 || p->lts > kMaxLevels) return fail(P, FS_E_ARG, "lts must be 0..%d", kMaxLevels);
  return FS_OK;
}

int ensure_rec_capacity(fs_pool *P, int n) {
  if (n <= P->capSlots) return FS_OK;
  int cap = std::max(16, P->capSlots * 2);
  while (cap < n) cap *= 2;
  StepRec *d = nullptr;
  HIPCHK(P, hipMalloc(&d, sizeof(StepRec) * cap));
  HIPCHK(P, hipMemsetAsync(d, 0, sizeof(StepRec) * cap, P->stream));
  if (P->drec) {
    HIPCHK(P, hipMemcpyAsync(d, P->drec, sizeof(StepRec) * P->capSlots, hipMemcpyDeviceToDevice, P->stream));
    HIPCHK(P, hipStreamSynchronize(P->stream));
    HIPCHK(P, hipFree(P->drec));
  }
  P->drec = d;
  P->capSlots = cap;
  P->grp.slots.clear();  // descriptors point into the old array
  return FS_OK;
}

void drop_graph(Group &G) {
  for (auto &g : G.graphs)
    if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
}

int upload_rec(fs_pool *P, int slot) {
  HIPCHK(P, hipMemcpyAsync(P->drec + slot, &P->slots[slot].hrec, sizeof(StepRec), hipMemcpyHostToDevice, P->stream));
  return FS_OK;
}

int download_recs(fs_pool *P, const std::vector<int> &slots) {
  if (slots.empty()) return FS_OK;
  int lo = *std::min_element(slots.begin(), slots.end()), hi = *std::max_element(slots.begin(), slots.end());
  std::vector<StepRec> tmp(hi - lo + 1);
  HIPCHK(P, hipMemcpyAsync(tmp.data(), P->drec + lo, sizeof(StepRec) * tmp.size(), hipMemcpyDeviceToHost, P->stream));
  HIPCHK(P, hipStreamSynchronize(P->stream));
  for (int s : slots) P->slots[s].hrec = tmp[s - lo];
  return FS_OK;
}

// the device descriptor of one slot whose strips are work items [wbeg, wend) of its group's list
DevInst devinst_of(const Slot &S, StepRec *rec, int wbeg, int wend) {
  DevInst d{};
  d.h = S.h; d.hu = S.hu; d.hv = S.hv; d.w = S.w;
  d.h1 = S.h1; d.hu1 = S.hu1; d.hv1 = S.hv1; d.w1 = S.w1;
  d.ab = (int)(S.h1 - S.h);
  d.z = S.z; d.n = S.n; d.mask = S.mask; d.hmax = S.hmax; d.infil = S.infil; d.rf = S.rf;
  d.rec = rec;
  d.nx = S.nx; d.ny = S.ny; d.pitch = S.pitch;
  d.flags = (S.infil ? FLAG_INFIL : 0) | (S.rf ? FLAG_RF : 0) | (S.p.implicitFriction ? FLAG_IMPF : 0) |
            (S.p.stepFlux && !S.L ? FLAG_STEPF : 0) |   // the step flux applies under global stepping only
            (S.stageCells ? FLAG_STAGE : 0);
  d.dx = (float)S.p.dx; d.inv_dx = (float)(1.0 / S.p.dx); d.g = (float)S.p.g; d.hdry = (float)S.p.hdry;
  d.hfront = (float)S.p.hfront; d.cfl = (float)S.p.cfl; d.dtDry = (float)S.p.dtDry;
  d.area = (float)(S.p.dx * S.p.dx);
  d.secondOrder = S.p.secondOrder ? 1 : 0;
  d.limiter = S.p.limiter;
  for (int k = 0; k < 4; ++k) d.bc[k] = S.p.bc[k];
  d.wbeg = wbeg;
  d.wend = wend;
  d.inl = S.inl; d.ispeed = S.ispeed; d.iapp = S.iapp; d.nin = S.nin;
  d.hl = S.L ? 4 : kHalo;
  d.own = strip_own(&S.p);
  d.rows = std::max(1, S.rows);
  d.sxn = S.sxn;
  // GI stores drain every step, dry or not: an instance with them steps every strip
  d.wet = S.nstrips && !S.p.untiled && !S.gi ? S.swet : nullptr;
  d.act = S.sact; d.srain = S.srain; d.sopen = S.sopen;
  d.hzA = S.hzA; d.hzB = S.hzB;
  d.L = S.L; d.tx = S.tx; d.ty = S.ty;
  d.lv = S.lv; d.lv0 = S.lv0; d.tsp = S.tsp; d.speedSlots = S.speedSlots; d.tileOpen = S.tileOpen;
  d.tileRf = S.rf ? S.tileRf : nullptr;
  d.st4 = S.st4;
  d.g4 = S.g4;
  d.gflag = S.gflag; d.bsx = S.bsx; d.bsy = S.bsy;
  d.urows = S.urows; d.ucells = S.ucells; d.plist = S.plist; d.capParts = kLtsUnitRuns * S.nunits;
  d.rend = S.rend; d.ticket = S.ticket;
  d.crest = S.crest; d.bidx = S.bidx; d.barFace = S.barFace; d.barGround = S.barGround;
  d.barFlux = S.barFlux; d.barBook = S.barBook; d.nbar = S.nbar;
  d.soil = S.soil; d.gi = S.gi;
  return d;
}

// extended faces and edges: the step flux, barrier crests, stage cells, soil, GI, the van Leer limiter or
// implicit friction (one kernel variant; the rule of ext_faces in the kernels)
bool ext_kind(const Slot &S, const DevInst &d) {
  return (d.flags & (FLAG_IMPF | FLAG_STEPF | FLAG_STAGE)) || S.nbar || S.soil || S.gi || S.p.limiter;
}

// (re)build the device descriptors and work list of a group of slots
int build_group(fs_pool *P, const std::vector<int> &slots) {
  Group &G = P->grp;
  if (G.slots == slots && G.dinst) return FS_OK;
  drop_graph(G);
  G.slots = slots;
  G.hinst.clear();
  G.withInlets.clear();
  G.withBarriers.clear();
  G.anyFirstOrder = G.anySecondOrder = false;
  G.anyHazard = false;
  G.anyStepFlux = false;
  G.kinds = 0;
  G.lts = false;
  std::vector<int4> work;
  for (size_t gi = 0; gi < slots.size(); ++gi) {
    Slot &S = P->slots[slots[gi]];
    const int wbeg = (int)work.size();
    const int rows = std::max(1, S.rows), own = strip_own(&S.p);
    for (int y0 = 0; y0 < S.ny; y0 += rows)
      for (int x0 = 0; x0 < S.nx; x0 += own) work.push_back(make_int4((int)gi, x0, y0, rows));
    const DevInst d = devinst_of(S, P->drec + slots[gi], wbeg, (int)work.size());
    const bool ext = ext_kind(S, d);
    if (ext) G.anyStepFlux = true;
    G.kinds |= 1u << ((S.hzA ? 2 : 0) + (ext ? 1 : 0));
    if (S.nbar) G.withBarriers.push_back((int)gi);
    if (S.L) G.lts = true;
    if (S.hzA) G.anyHazard = true;
    if (S.nin) G.withInlets.push_back((int)gi);
    if (!S.p.secondOrder) G.anyFirstOrder = true;
    else G.anySecondOrder = true;
    G.hinst.push_back(d);
  }
  const int ni = (int)G.hinst.size(), nw = (int)work.size();
  if (G.lts && ni > 1) { G.slots.clear(); return fail(P, FS_E_ARG, "an instance with local time stepping advances alone"); }
  if (ni > G.capInst) {
    if (G.dinst) HIPCHK(P, hipFree(G.dinst));
    G.capInst = std::max(ni, G.capInst * 2);
    HIPCHK(P, hipMalloc(&G.dinst, sizeof(DevInst) * G.capInst));
  }
  // partials: one per wavefront (an LTS record launches up to kLtsWaveCap per unit)
  const int needWork = G.lts ? kLtsWaveCap * nw : nw;
  if (needWork > G.capWork) {
    if (G.dwork) HIPCHK(P, hipFree(G.dwork));
    if (G.p1) HIPCHK(P, hipFree(G.p1));
    if (G.p2) HIPCHK(P, hipFree(G.p2));
    G.capWork = std::max(needWork, G.capWork * 2);
    HIPCHK(P, hipMalloc(&G.dwork, sizeof(int4) * G.capWork));
    HIPCHK(P, hipMalloc(&G.p1, sizeof(Part) * G.capWork));
    HIPCHK(P, hipMalloc(&G.p2, sizeof(Part) * G.capWork));
  }
  if (ni) HIPCHK(P, hipMemcpyAsync(G.dinst, G.hinst.data(), sizeof(DevInst) * ni, hipMemcpyHostToDevice, P->stream));
  if (nw) HIPCHK(P, hipMemcpyAsync(G.dwork, work.data(), sizeof(int4) * nw, hipMemcpyHostToDevice, P->stream));
  G.nwork = nw;
  return FS_OK;
}

__global__ void sw_set_targets(const DevInst *__restrict__ insts, int n, const double *__restrict__ t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) insts[i].rec->tEnd = t[i];
}

// ---- the allocation-free advance (fs_advance_noalloc): descriptor, work list and per-wavefront partials of
// one solver in a file-scope device array of fixed size, written by kernels from launch arguments (no host
// buffers, no allocation). The array is the process's for its lifetime, one copy per device; its device
// address comes from hipGetSymbolAddress on the first call on each device.
constexpr size_t kNoAllocScratch = (size_t)4 << 20;
__device__ __align__(256) char scratch[kNoAllocScratch];

__global__ void sw_put_inst(DevInst *__restrict__ dst, DevInst d, double tEnd) {
  *dst = d;
  d.rec->tEnd = tEnd;
}

// the work list of one instance, in build_group's order: strips across (x fastest), then down
__global__ __launch_bounds__(256) void sw_fill_work(int4 *__restrict__ work, int n, int sxn, int own, int rows) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < n) work[w] = make_int4(0, (w % sxn) * own, (w / sxn) * rows, rows);
}

hipEvent_t span_event(fs_pool *P) {
  if (P->nSpanEv >= (int)P->spanEv.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    P->spanEv.push_back(e);
  }
  return P->spanEv[P->nSpanEv++];
}

hipEvent_t timing_event(fs_pool *P) {
  if (P->nTev >= (int)P->tev.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    P->tev.push_back(e);
  }
  return P->tev[P->nTev++];
}

int launch_step(fs_pool *P, const Launch &G) {
  const int nb = (G.nwork + kGlobalWaves - 1) / kGlobalWaves, ni = G.ni;
  constexpr int kT = 64 * kGlobalWaves;
  hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
  if (P->timing) {
    e0 = timing_event(P); e1 = timing_event(P); e2 = timing_event(P);
    if (!e0 || !e1 || !e2) return fail(P, FS_E_HIP, "hipEventCreate failed");
    HIPCHK(P, hipEventRecord(e0, P->stream));
  }
  // one launch per kernel variant the group's instances need (each variant skips the other kinds)
  const unsigned K = G.kinds;
  if (K & 0x5u) sw_stage<0, false><<<nb, kT, 0, P->stream>>>(G.dinst, G.dwork, G.nwork, G.p1, 0);
  if (K & 0xau) sw_stage<0, false, false, true><<<nb, kT, 0, P->stream>>>(G.dinst, G.dwork, G.nwork, G.p1, 0);
  for (int q = 0; q < G.nWithBarriers; ++q) {
    const int gi = G.withBarriers[q];
    sw_barrier_books<0, false><<<(G.hinst[gi].nbar + 63) / 64, 64, 0, P->stream>>>(G.dinst, gi, 0);
  }
  if (P->timing) HIPCHK(P, hipEventRecord(e1, P->stream));
  // each instance runs its own scheme: the kernel of the other order returns at once for it
  if (G.anySecondOrder) {
    if (K & 0x1u) sw_stage<1, false><<<nb, kT, 0, P->stream>>>(G.dinst, G.dwork, G.nwork, G.p2, 0);
    if (K & 0x2u) sw_stage<1, false, false, true><<<nb, kT, 0, P->stream>>>(G.dinst, G.dwork, G.nwork, G.p2, 0);
    if (K & 0x4u) sw_stage<1, false, true><<<nb, kT, 0, P->stream>>>(G.dinst, G.dwork, G.nwork, G.p2, 0);
    if (K & 0x8u) sw_stage<1, false, true, true><<<nb, kT, 0, P->stream>>>(G.dinst, G.dwork, G.nwork, G.p2, 0);
    for (int q = 0; q < G.nWithBarriers; ++q) {
      const int gi = G.withBarriers[q];
      sw_barrier_books<1, false><<<(G.hinst[gi].nbar + 63) / 64, 64, 0, P->stream>>>(G.dinst, gi, 0);
    }
  }
  if (P->timing) P->spans.push_back({e1, nullptr, -1, -1});  // completed below with e2
  if (G.anyFirstOrder) {
    if (K & 0x1u) sw_euler_finish<false><<<nb, kT, 0, P->stream>>>(G.dinst, G.dwork, G.nwork, G.p2, 0);
    if (K & 0x2u) sw_euler_finish<false, false, true><<<nb, kT, 0, P->stream>>>(G.dinst, G.dwork, G.nwork, G.p2, 0);
    if (K & 0x4u) sw_euler_finish<false, true><<<nb, kT, 0, P->stream>>>(G.dinst, G.dwork, G.nwork, G.p2, 0);
    if (K & 0x8u) sw_euler_finish<false, true, true><<<nb, kT, 0, P->stream>>>(G.dinst, G.dwork, G.nwork, G.p2, 0);
  }
  if (P->timing) { HIPCHK(P, hipEventRecord(e2, P->stream)); P->spans.back().b = e2; }
  for (int q = 0; q < G.nWithInlets; ++q) {
    const int gi = G.withInlets[q];
    const int nin = G.hinst[gi].nin;
    sw_exchange<false><<<(nin + 63) / 64, 64, 0, P->stream>>>(G.dinst, gi, 0);
  }
  sw_dt_final<<<ni, 1024, 0, P->stream>>>(G.dinst, G.p1, G.p2, 0);
  HIPCHK(P, hipGetLastError());
  return FS_OK;
}

// fold the recorded event triples into the stage totals
int collect_timing(fs_pool *P) {
  if (!P->nTev) return FS_OK;
  HIPCHK(P, hipEventSynchronize(P->tev[P->nTev - 1]));
  for (int i = 0; i + 2 < P->nTev; i += 3) {
    float a = 0.f, b = 0.f;
    HIPCHK(P, hipEventElapsedTime(&a, P->tev[i], P->tev[i + 1]));
    HIPCHK(P, hipEventElapsedTime(&b, P->tev[i + 1], P->tev[i + 2]));
    P->ms1 += a;
    P->ms2 += b;
    P->nTimed += 1;
  }
  P->nTev = 0;
  if (!P->spans.empty()) {
    HIPCHK(P, hipEventSynchronize(P->spans.back().b));
    std::vector<int32_t> hist;
    if (P->nCycles) {
      hist.resize((size_t)(kMaxLevels + 1) * P->nCycles);
      HIPCHK(P, hipMemcpy(hist.data(), P->dhist, sizeof(int32_t) * hist.size(), hipMemcpyDeviceToHost));
    }
    double open = 0.0;
    for (int s : P->grp.slots) open += P->slots[s].openCells;
    for (auto &sp : P->spans) {
      float ms = 0.f;
      HIPCHK(P, hipEventElapsedTime(&ms, sp.a, sp.b));
      P->domMs += ms;
      P->domLaunches += 1;
      P->domCells += sp.cycle < 0 ? open : (double)hist[(size_t)(kMaxLevels + 1) * sp.cycle + sp.level];
    }
    P->spans.clear();
    P->nSpanEv = 0;
    P->nCycles = 0;
  }
  return FS_OK;
}

// one local-time-stepping cycle of the group's (single) instance: levels and lists, then the
// records finest first (solver-cpu.js:504-527; solver-gpu.js:367-396)
int launch_cycle(fs_pool *P, const Launch &G) {
  const DevInst &d = G.hinst[0];
  const int nt = d.tx * d.ty, nu = d.wend - d.wbeg, L = d.L;
  // kLtsWaveCap wavefronts per unit (a record's parts beyond that share wavefronts); idle blocks exit
  const int nbp = ((kLtsWaveCap * nu + kLtsWaves - 1) / kLtsWaves + 7) / 8 * 8;
  constexpr int kT = 64 * kLtsWaves;
  hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
  if (P->timing) {
    e0 = timing_event(P); e1 = timing_event(P); e2 = timing_event(P);
    if (!e0 || !e1 || !e2) return fail(P, FS_E_HIP, "hipEventCreate failed");
    HIPCHK(P, hipEventRecord(e0, P->stream));
  }
  sw_lts_speed<<<((d.tx + 7) / 8 * d.ty + 3) / 4, 256, 0, P->stream>>>(G.dinst);
  sw_lts_cycle<<<1, 1024, 0, P->stream>>>(G.dinst);
  sw_lts_level<<<(nt + 255) / 256, 256, 0, P->stream>>>(G.dinst);
  sw_lts_smooth<<<(nt + 255) / 256, 256, 0, P->stream>>>(G.dinst);
  sw_lts_units<<<(nu + 255) / 256, 256, 0, P->stream>>>(G.dinst, G.dwork);
  sw_lts_lists<<<kMaxLevels + 1, 1024, 0, P->stream>>>(G.dinst, G.dwork);
  const int cyc = P->nCycles;
  if (P->timing) {
    HIPCHK(P, hipEventRecord(e1, P->stream));
    if (cyc >= P->capHist) {
      const int cap = std::max(64, 2 * P->capHist);
      int32_t *h = nullptr;
      HIPCHK(P, hipStreamSynchronize(P->stream));
      HIPCHK(P, hipMalloc(&h, sizeof(int32_t) * (kMaxLevels + 1) * cap));
      if (P->dhist) {
        HIPCHK(P, hipMemcpy(h, P->dhist, sizeof(int32_t) * (kMaxLevels + 1) * P->capHist, hipMemcpyDeviceToDevice));
        HIPCHK(P, hipFree(P->dhist));
      }
      P->dhist = h;
      P->capHist = cap;
    }
    HIPCHK(P, hipMemcpyAsync(P->dhist + (size_t)(kMaxLevels + 1) * cyc, &P->drec[G.slot0].lcells,
                             sizeof(int32_t) * (kMaxLevels + 1), hipMemcpyDeviceToDevice, P->stream));
    P->nCycles += 1;
  }
  const int Pn = 1 << L;
  for (int k = 0; k < Pn; ++k)
    for (int m = 0; m <= L; ++m) {
      if ((k + 1) % (1 << m)) continue;
      const bool ext = d.nbar || (d.flags & FLAG_STAGE) || d.soil || d.gi || d.limiter;
      if (ext) {
        sw_stage<0, true, false, true><<<nbp, kT, 0, P->stream>>>(G.dinst, G.dwork, G.nwork, G.p1, m);
        if (d.nbar) sw_barrier_books<0, true><<<(d.nbar + 63) / 64, 64, 0, P->stream>>>(G.dinst, 0, m);
      } else {
        sw_stage<0, true><<<nbp, kT, 0, P->stream>>>(G.dinst, G.dwork, G.nwork, G.p1, m);
      }
      hipEvent_t sa = nullptr, sb = nullptr;
      if (P->timing) {
        sa = span_event(P); sb = span_event(P);
        if (!sa || !sb) return fail(P, FS_E_HIP, "hipEventCreate failed");
        HIPCHK(P, hipEventRecord(sa, P->stream));
      }
      // the record's base step rides in the level argument's upper bits (its end time: hazard clock)
      const int mk = m | (k << 8);
      if (d.secondOrder) {
        if (ext) {
          if (d.hzA) sw_stage<1, true, true, true><<<nbp, kT, 0, P->stream>>>(G.dinst, G.dwork, G.nwork, G.p2, mk);
          else sw_stage<1, true, false, true><<<nbp, kT, 0, P->stream>>>(G.dinst, G.dwork, G.nwork, G.p2, mk);
          if (d.nbar) sw_barrier_books<1, true><<<(d.nbar + 63) / 64, 64, 0, P->stream>>>(G.dinst, 0, m);
        } else if (d.hzA) {
          sw_stage<1, true, true><<<nbp, kT, 0, P->stream>>>(G.dinst, G.dwork, G.nwork, G.p2, mk);
        } else {
          sw_stage<1, true><<<nbp, kT, 0, P->stream>>>(G.dinst, G.dwork, G.nwork, G.p2, mk);
        }
      } else {
        if (ext) {
          if (d.hzA) sw_euler_finish<true, true, true><<<nbp, kT, 0, P->stream>>>(G.dinst, G.dwork, G.nwork, G.p2, mk);
          else sw_euler_finish<true, false, true><<<nbp, kT, 0, P->stream>>>(G.dinst, G.dwork, G.nwork, G.p2, mk);
        } else if (d.hzA) {
          sw_euler_finish<true, true><<<nbp, kT, 0, P->stream>>>(G.dinst, G.dwork, G.nwork, G.p2, mk);
        } else {
          sw_euler_finish<true><<<nbp, kT, 0, P->stream>>>(G.dinst, G.dwork, G.nwork, G.p2, mk);
        }
      }
      if (P->timing) {
        HIPCHK(P, hipEventRecord(sb, P->stream));
        P->spans.push_back({sa, sb, cyc, m});
      }
      // books and inlet exchange (first order: the gathers added first, by every part's wavefront)
      sw_lts_fold<<<d.secondOrder ? kLtsEndSlices : nbp, kT, 0, P->stream>>>(G.dinst, G.p1, G.p2, m);
    }
  sw_lts_cycle_end<<<1, 1, 0, P->stream>>>(G.dinst);
  if (P->timing) HIPCHK(P, hipEventRecord(e2, P->stream));
  HIPCHK(P, hipGetLastError());
  return FS_OK;
}

int launch_steps(fs_pool *P, int K) {
  Group &G = P->grp;
  const Launch L = view_of(G);
  auto one = [&]() { return L.lts ? launch_cycle(P, L) : launch_step(P, L); };
  if (P->timing || !P->useGraphs || K < 2) {
    for (int k = 0; k < K; ++k) {
      int rc = one();
      if (rc) return rc;
    }
    return P->timing && (P->nTev > 3 * 256 || P->spans.size() > 4096) ? collect_timing(P) : FS_OK;
  }
  // a batch is a run of captured power-of-two graphs (K = 64 + 8: the 64- and the 8-step graph)
  while (K > 0) {
    int b = 0;
    while (b + 1 < 16 && (2 << b) <= K) ++b;
    hipGraphExec_t &ge = G.graphs[b];
    if (!ge) {
      hipGraph_t g = nullptr;
      HIPCHK(P, hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal));
      int rc = FS_OK;
      for (int k = 0; k < (1 << b) && !rc; ++k) rc = one();
      hipError_t e = hipStreamEndCapture(P->stream, &g);
      if (rc) return rc;
      if (e != hipSuccess) return fail(P, FS_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
      e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) { ge = nullptr; return fail(P, FS_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e)); }
    }
    HIPCHK(P, hipGraphLaunch(ge, P->stream));
    K -= 1 << b;
  }
  return FS_OK;
}

// start of a launch group: targets, then a fresh dt (dirty instances) or a re-clamped one
int begin_advance(fs_pool *P, const std::vector<int> &slots, const std::vector<double> &targets) {
  int rc = build_group(P, slots);
  if (rc) return rc;
  Group &G = P->grp;
  const int ni = (int)slots.size();
  if (ni > P->capTarget) {
    if (P->dtarget) HIPCHK(P, hipFree(P->dtarget));
    P->capTarget = std::max(ni, 2 * P->capTarget);
    HIPCHK(P, hipMalloc(&P->dtarget, sizeof(double) * P->capTarget));
  }
  HIPCHK(P, hipMemcpyAsync(P->dtarget, targets.data(), sizeof(double) * ni, hipMemcpyHostToDevice, P->stream));
  sw_set_targets<<<(ni + 63) / 64, 64, 0, P->stream>>>(G.dinst, ni, P->dtarget);
  bool dirty = false;
  for (int s : slots) dirty = dirty || P->slots[s].hrec.dirty;
  if (!G.lts) {
    if (dirty) sw_speed<<<(G.nwork + 3) / 4, 256, 0, P->stream>>>(G.dinst, G.dwork, G.nwork, G.p2);
    sw_dt_final<<<ni, 1024, 0, P->stream>>>(G.dinst, G.p1, G.p2, 1);
  }
  HIPCHK(P, hipGetLastError());
  for (int s : slots) P->slots[s].hrec.dirty = 0;
  return FS_OK;
}

void fill_clock(const Slot &S, fs_clock *c) {
  c->t = S.hrec.t;
  c->lastDt = S.hrec.lastDt;
  c->dtAvg = S.dtSteps ? S.dtSum / (double)S.dtSteps : 0.0;
  c->steps = S.hrec.steps;
  c->cellUpdates = (int64_t)S.hrec.cellUpdates;
}

int valid_slot(fs_pool *P, int slot) {
  if (!P) return FS_E_ARG;
  if (slot < 0 || slot >= (int)P->slots.size() || !P->slots[slot].used) return fail(P, FS_E_ARG, "bad slot %d", slot);
  return FS_OK;
}

int advance_group(fs_pool *P, const int32_t *slots, const double *targets, int n, int K, bool once) {
  if (n <= 0) return FS_OK;
  if (K < 1) K = 1;
  std::vector<int> sl(slots, slots + n);
  std::vector<double> tg(targets, targets + n);
  for (int i = 0; i < n; ++i) {
    int rc = valid_slot(P, sl[i]);
    if (rc) return rc;
  }
  // sorted unique slot order keeps the group (and its graph) stable across calls
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int a, int b) { return sl[a] < sl[b]; });
  std::vector<int> gs;
  std::vector<double> gt;
  for (int i : order) {
    if (!gs.empty() && gs.back() == sl[i]) return fail(P, FS_E_ARG, "slot %d listed twice", sl[i]);
    gs.push_back(sl[i]);
    gt.push_back(tg[i]);
  }
  int rc = begin_advance(P, gs, gt);
  if (rc) return rc;
  std::vector<double> tprev(gs.size());
  for (size_t i = 0; i < gs.size(); ++i) tprev[i] = P->slots[gs[i]].hrec.t;
  // one launch sequence holds at most kMaxBatch steps (a captured graph is sized by it); a frame
  // (once) runs at most K steps in total, an advance repeats batches until every target is reached
  constexpr int kMaxBatch = 256;
  int64_t left = K;
  int stalls = 0;
  std::vector<double> recent(gs.size(), 0.0);   // mean step of each slot's last batch
  for (;;) {
    int kb = (int)std::min<int64_t>(left, kMaxBatch);
    if (!once) {
      // size an advance's batch to the steps still needed at the mean step so far, so that few
      // zero-length steps run past the targets (the first batch, with no mean yet, takes K)
      double need = 0.0;
      bool known = true;
      for (size_t i = 0; i < gs.size(); ++i) {
        const Slot &S = P->slots[gs[i]];
        const double mean = recent[i] > 0.0 ? recent[i] : (S.dtSteps > 0 ? S.dtSum / (double)S.dtSteps : 0.0);
        if (mean > 0.0) need = std::max(need, (gt[i] - S.hrec.t) / mean);
        else known = false;
      }
      if (known) kb = (int)std::max<double>(1.0, std::min<double>(kb, std::ceil(need * 1.02) + 1.0));
    }
    rc = P->noSteps ? FS_OK : launch_steps(P, kb);
    if (rc) return rc;
    if (once) left -= kb;
    rc = download_recs(P, gs);
    if (rc) return rc;
    bool done = true, progress = false;
    for (size_t i = 0; i < gs.size(); ++i) {
      Slot &S = P->slots[gs[i]];
      if (S.hrec.t > tprev[i]) {
        progress = true;
        const int64_t ds = S.hrec.steps - S.dtSteps;
        if (ds > 0) {
          S.dtSum += S.hrec.t - tprev[i];
          S.dtSteps = S.hrec.steps;
          recent[i] = (S.hrec.t - tprev[i]) / (double)ds;
        }
      }
      tprev[i] = S.hrec.t;
      if (S.hrec.t < gt[i]) done = false;
    }
    if (done || (once && left <= 0)) break;
    if (!progress && ++stalls > 3) return fail(P, FS_E_STALL, "GPU solver made no progress (device error?)");
  }
  return FS_OK;
}

// the device address of the scratch array on the current device (hipGetSymbolAddress once per device)
int scratch_base(fs_pool *P, char **out) {
  static std::mutex mu;
  static char *base[64] = {};
  int dev = -1;
  HIPCHK(P, hipGetDevice(&dev));
  if (dev != P->device)
    return fail(P, FS_E_STATE, "fs_advance_noalloc: device %d is current, the solver lives on device %d", dev, P->device);
  if (dev < 0 || dev >= 64) return fail(P, FS_E_STATE, "fs_advance_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(mu);
  if (!base[dev]) {
    char *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(P, FS_E_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    base[dev] = ptr;
  }
  *out = base[dev];
  return FS_OK;
}

// fs_advance without allocation: the launch sequence of fs_advance (the same kernels in the same order, plain
// launches instead of captured graphs) over the scratch array; one slot
int advance_noalloc(fs_pool *P, int slot, double target, int K, fs_clock *clk) {
  int rc = valid_slot(P, slot);
  if (rc) return rc;
  if (P->timing) return fail(P, FS_E_STATE, "fs_advance_noalloc: timing mode creates HIP events; switch it off");
  Slot &S = P->slots[slot];
  const int own = strip_own(&S.p), rows = std::max(1, S.rows);
  const int sxn = (S.nx + own - 1) / own, nwork = sxn * ((S.ny + rows - 1) / rows);
  const int needWork = S.L ? kLtsWaveCap * nwork : nwork;
  const size_t bInst = align256(sizeof(DevInst)), bWork = align256(sizeof(int4) * (size_t)nwork);
  const size_t bPart = align256(sizeof(Part) * (size_t)needWork), need = bInst + bWork + 2 * bPart;
  if (need > kNoAllocScratch)
    return fail(P, FS_E_ARG, "fs_advance_noalloc: this grid needs %zu bytes of scratch (%d work items of %d x %d "
                "cells), more than the %zu bytes the library holds", need, nwork, own, rows, kNoAllocScratch);
  char *base = nullptr;
  if ((rc = scratch_base(P, &base))) return rc;
  if (K < 1) K = 1;
  const DevInst d = devinst_of(S, P->drec + slot, 0, nwork);
  const int zero = 0;
  Launch L;
  L.dinst = (DevInst *)base;
  L.dwork = (int4 *)(base + bInst);
  L.p1 = (Part *)(base + bInst + bWork);
  L.p2 = (Part *)(base + bInst + bWork + bPart);
  L.nwork = nwork; L.ni = 1;
  L.hinst = &d;
  L.withInlets = &zero; L.nWithInlets = S.nin ? 1 : 0;
  L.withBarriers = &zero; L.nWithBarriers = S.nbar ? 1 : 0;
  L.kinds = 1u << ((S.hzA ? 2 : 0) + (ext_kind(S, d) ? 1 : 0));
  L.anyFirstOrder = !S.p.secondOrder; L.anySecondOrder = S.p.secondOrder != 0; L.lts = S.L > 0;
  L.slot0 = slot;
  // descriptor and target, work list; then (global stepping) the first dt, as begin_advance
  sw_put_inst<<<1, 1, 0, P->stream>>>(L.dinst, d, target);
  sw_fill_work<<<(nwork + 255) / 256, 256, 0, P->stream>>>(L.dwork, nwork, sxn, own, rows);
  if (!L.lts) {
    if (S.hrec.dirty) sw_speed<<<(nwork + 3) / 4, 256, 0, P->stream>>>(L.dinst, L.dwork, nwork, L.p2);
    sw_dt_final<<<1, 1024, 0, P->stream>>>(L.dinst, L.p1, L.p2, 1);
  }
  HIPCHK(P, hipGetLastError());
  S.hrec.dirty = 0;
  // batches as advance_group: sized to the steps still needed at the recent mean step
  constexpr int kMaxBatch = 256;
  double tprev = S.hrec.t, recent = 0.0;
  int stalls = 0;
  for (;;) {
    int kb = K < kMaxBatch ? K : kMaxBatch;
    const double mean = recent > 0.0 ? recent : (S.dtSteps > 0 ? S.dtSum / (double)S.dtSteps : 0.0);
    if (mean > 0.0) kb = (int)std::max<double>(1.0, std::min<double>(kb, std::ceil((target - S.hrec.t) / mean * 1.02) + 1.0));
    for (int k = 0; k < kb && !P->noSteps; ++k)
      if ((rc = L.lts ? launch_cycle(P, L) : launch_step(P, L))) return rc;
    StepRec r;
    HIPCHK(P, hipMemcpyAsync(&r, P->drec + slot, sizeof r, hipMemcpyDeviceToHost, P->stream));
    HIPCHK(P, hipStreamSynchronize(P->stream));
    S.hrec = r;
    bool progress = false;
    if (S.hrec.t > tprev) {
      progress = true;
      const int64_t ds = S.hrec.steps - S.dtSteps;
      if (ds > 0) {
        S.dtSum += S.hrec.t - tprev;
        S.dtSteps = S.hrec.steps;
        recent = (S.hrec.t - tprev) / (double)ds;
      }
    }
    tprev = S.hrec.t;
    if (S.hrec.t >= target) break;
    if (!progress && ++stalls > 3) return fail(P, FS_E_STALL, "GPU solver made no progress (device error?)");
  }
  if (clk) fill_clock(S, clk);
  return FS_OK;
}

// strip of interior cell (i, j) under global stepping (work items run x fastest)
inline int strip_of(const Slot &S, int i, int j) { return (j / S.rows) * S.sxn + i / strip_own(&S.p); }

// per strip: bit 0 an open cell with a positive rain multiplier (all open cells without a rain field),
// bit 1 a stage cell (such strips always step: the stage can flood a dry edge)
int upload_strip_rain(fs_pool *P, Slot &S) {
  if (!S.nstrips) return FS_OK;
  std::vector<uint8_t> sr(S.nstrips, 0);
  for (int j = 0; j < S.ny; ++j)
    for (int i = 0; i < S.nx; ++i) {
      const size_t k = (size_t)j * S.nx + i;
      if ((S.hmask[k] & MASK_OPEN) && (S.hrf.empty() || S.hrf[k] > 0.f)) sr[strip_of(S, i, j)] |= 1;
      if (S.hmask[k] & MASK_STAGE) sr[strip_of(S, i, j)] |= 2;
    }
  HIPCHK(P, hipMemcpy(S.srain, sr.data(), (size_t)S.nstrips, hipMemcpyHostToDevice));
  return FS_OK;
}

// LTS: open cells and rain multiplier (the host copy of the rain field; 1 without one) summed per 8x8 tile,
// for the per-level books of each cycle
int upload_tile_tables(fs_pool *P, Slot &S) {
  const size_t nt = (size_t)S.tx * S.ty;
  std::vector<int32_t> to(nt, 0);
  std::vector<float> tr(nt, 0.f);
  for (int j = 0; j < S.ny; ++j)
    for (int i = 0; i < S.nx; ++i) {
      const size_t k = (size_t)j * S.nx + i;
      if (S.hmask[k] & MASK_OPEN) {
        to[(size_t)(j / 8) * S.tx + i / 8] += 1;
        tr[(size_t)(j / 8) * S.tx + i / 8] += S.hrf.empty() ? 1.f : S.hrf[k];
      }
    }
  HIPCHK(P, hipMemcpy(S.tileOpen, to.data(), nt * 4, hipMemcpyHostToDevice));
  HIPCHK(P, hipMemcpy(S.tileRf, tr.data(), nt * 4, hipMemcpyHostToDevice));
  return FS_OK;
}

// hazard products before the first step: nothing recorded, no arrival, no first wetting
__global__ __launch_bounds__(256) void sw_hz_init(float4 *__restrict__ a, float4 *__restrict__ b, size_t n) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
    a[k] = make_float4(0.f, 0.f, 0.f, -1.f);
    b[k] = make_float4(0.f, -1.f, 0.f, 0.f);
  }
}

// per strip: open cells, rain flag, and every strip marked as stepping (its U* is then cleared by
// the first stage that finds it quiet)
int upload_strip_tables(fs_pool *P, Slot &S) {
  std::vector<int32_t> so(S.nstrips, 0);
  for (int j = 0; j < S.ny; ++j)
    for (int i = 0; i < S.nx; ++i)
      if (S.hmask[(size_t)j * S.nx + i] & MASK_OPEN) so[strip_of(S, i, j)] += 1;
  HIPCHK(P, hipMemcpy(S.sopen, so.data(), (size_t)S.nstrips * 4, hipMemcpyHostToDevice));
  HIPCHK(P, hipMemset(S.sact, 1, (size_t)S.nstrips));
  return upload_strip_rain(P, S);
}

// per-cell host preparation shared by create/add
int add_slot(fs_pool *P, const fs_params *p, const double *z, const double *manning, double manningScalar,
             const uint8_t *mask, const double *infil, const double *rainField, const int32_t *inlets, int nInlets,
             int32_t *slotOut) {
  int rc = check_params(P, p);
  if (rc) return rc;
  if (!z) return fail(P, FS_E_ARG, "z is NULL");
  if (nInlets < 0 || (nInlets > 0 && !inlets)) return fail(P, FS_E_ARG, "bad inlets");
  const int nx = p->nx, ny = p->ny;
  const size_t N = (size_t)nx * ny;
  for (int m = 0; m < nInlets; ++m)
    if (inlets[m] < 0 || (size_t)inlets[m] >= N) return fail(P, FS_E_ARG, "inlet %d out of the grid", m);
  const size_t need = slot_bytes(p, infil != nullptr, rainField != nullptr, nInlets);
  if (P->budget && P->used + need > P->budget)
    return fail(P, FS_E_BUDGET, "HBM budget: %zu bytes in use, %zu needed, budget %zu", P->used, need, P->budget);
  int slot = -1;
  for (size_t s = 0; s < P->slots.size(); ++s)
    if (!P->slots[s].used) { slot = (int)s; break; }
  if (slot < 0) { slot = (int)P->slots.size(); P->slots.emplace_back(); }
  rc = ensure_rec_capacity(P, (int)P->slots.size());
  if (rc) return rc;
  Slot &S = P->slots[slot];
  S = Slot();
  S.p = *p;
  S.nx = nx; S.ny = ny; S.pitch = nx;
  S.rows = p->rows > 0 ? p->rows : 64;
  S.bytes = need;
  hipError_t e = hipMalloc(&S.mem, need);
  if (e != hipSuccess) return fail(P, FS_E_HIP, "hipMalloc(%zu): %s", need, hipGetErrorString(e));
  char *q = S.mem;
  auto take = [&](size_t b) { char *r = q; q += align256(b); return r; };
  // each variable's two states side by side (h, h1, hu, hu1, ...): a lane reads U^n or U* of a variable from one
  // plane base at a per-lane offset (the LTS second stage mixes them within a row)
  S.h = (float *)take(N * 4); S.h1 = (float *)take(N * 4); S.hu = (float *)take(N * 4); S.hu1 = (float *)take(N * 4);
  S.hv = (float *)take(N * 4); S.hv1 = (float *)take(N * 4); S.w = (float *)take(N * 4); S.w1 = (float *)take(N * 4);
  S.z = (float *)take(N * 4); S.n = (float *)take(N * 4); S.hmax = (float *)take(N * 4);
  S.mask = (uint8_t *)take(N);
  if (infil) S.infil = (float *)take(N * 4);
  if (rainField) S.rf = (float *)take(N * 4);
  if (p->hazard) { S.hzA = (float4 *)take(N * 16); S.hzB = (float4 *)take(N * 16); }
  S.nin = nInlets;
  S.inl = (Inlet *)take((size_t)std::max(nInlets, 1) * sizeof(Inlet));
  S.ispeed = (float *)take((size_t)std::max(nInlets, 1) * 4);
  S.iapp = (float4 *)take((size_t)std::max(nInlets, 1) * 16);
  S.rows = strip_rows(p);
  if (p->lts) {
    S.L = p->lts;
    S.tx = (nx + 7) / 8;
    S.ty = (ny + 7) / 8;
    const size_t nt = (size_t)S.tx * S.ty;
    S.nunits = units_of(p);
    S.lv = (uint8_t *)take(nt); S.lv0 = (uint8_t *)take(nt);
    S.tsp = (float *)take(nt * 4); S.tileOpen = (int32_t *)take(nt * 4); S.tileRf = (float *)take(nt * 4);
    S.speedSlots = (unsigned *)take(kSpeedSlots * 4);
    S.st4 = (float4 *)take(N * 16);
    S.g4 = (float4 *)take(N * 16);
    S.gflag = (uint8_t *)take(N);
    S.bsx = (float4 *)take(N * 16); S.bsy = (float4 *)take(N * 16);
    S.urows = (uint64_t *)take((size_t)S.nunits * 8 * (kMaxLevels + 1));
    S.ucells = (int32_t *)take((size_t)S.nunits * 4 * (kMaxLevels + 1));
    S.plist = (int4 *)take((size_t)S.nunits * kLtsUnitRuns * 16 * (kMaxLevels + 1));
    S.rend = (double *)take((size_t)kLtsEndSlices * 3 * 8);
    S.ticket = (unsigned *)take(4);
  } else {
    S.sxn = (nx + strip_own(p) - 1) / strip_own(p);
    S.nstrips = units_of(p);
    const size_t cap = strip_cap(p);
    S.swet = (uint8_t *)take(cap); S.sact = (uint8_t *)take(cap);
    S.srain = (uint8_t *)take(cap); S.sopen = (int32_t *)take(cap * 4);
  }
  const size_t carved = (size_t)(q - S.mem);
  if (carved > need) {   // the planes carved out must fit what slot_bytes counted
    (void)hipFree(S.mem);
    S.mem = nullptr;
    return fail(P, FS_E_ARG, "internal: the instance's planes take %zu bytes, %zu counted", carved, need);
  }
  S.used = true;
  P->used += need;

  double zmin = INFINITY;
  for (size_t k = 0; k < N; ++k) zmin = std::min(zmin, z[k]);
  S.zref = zmin;
  std::vector<float> tmp(N);
  for (size_t k = 0; k < N; ++k) tmp[k] = (float)(z[k] - zmin);
  HIPCHK(P, hipMemcpy(S.z, tmp.data(), N * 4, hipMemcpyHostToDevice));
  for (size_t k = 0; k < N; ++k) tmp[k] = (float)(manning ? manning[k] : manningScalar);
  HIPCHK(P, hipMemcpy(S.n, tmp.data(), N * 4, hipMemcpyHostToDevice));
  S.hmask.assign(N, 0);
  int open = 0;
  for (size_t k = 0; k < N; ++k) {
    const bool o = mask ? mask[k] != 0 : true;
    S.hmask[k] = (o ? MASK_OPEN : 0) | (tmp[k] > 0.f ? MASK_ROUGH : 0);   // tmp: the FP32 Manning n
    open += o;
  }
  S.inletCells.assign(inlets, inlets + nInlets);
  for (int m = 0; m < nInlets; ++m) S.hmask[inlets[m]] |= MASK_INLET;
  S.openCells = open;
  HIPCHK(P, hipMemcpy(S.mask, S.hmask.data(), N, hipMemcpyHostToDevice));
  if (infil) {
    for (size_t k = 0; k < N; ++k) tmp[k] = (float)infil[k];
    HIPCHK(P, hipMemcpy(S.infil, tmp.data(), N * 4, hipMemcpyHostToDevice));
  }
  if (rainField) {
    S.hrf.resize(N);
    for (size_t k = 0; k < N; ++k) S.hrf[k] = (float)rainField[k];
  }
  if (S.L) {
    const size_t nt = (size_t)S.tx * S.ty;
    rc = upload_tile_tables(P, S);
    if (rc) return rc;
    HIPCHK(P, hipMemset(S.lv, 0, nt));
    HIPCHK(P, hipMemset(S.speedSlots, 0, kSpeedSlots * 4));
    HIPCHK(P, hipMemset(S.st4, 0, (size_t)((char *)S.bsx - (char *)S.st4)));    // store, gathers, flags
    HIPCHK(P, hipMemset(S.bsx, 0, N * 16));   // each plane on its own: take() pads them apart
    HIPCHK(P, hipMemset(S.bsy, 0, N * 16));
    HIPCHK(P, hipMemset(S.ticket, 0, 4));     // (each record ends with it at 0 again)
    HIPCHK(P, hipDeviceSynchronize());
  }
  double rainArea = 0.0;
  if (rainField) {
    for (size_t k = 0; k < N; ++k) if (S.hmask[k] & MASK_OPEN) rainArea += S.hrf[k];
    HIPCHK(P, hipMemcpy(S.rf, S.hrf.data(), N * 4, hipMemcpyHostToDevice));
  } else rainArea = open;
  if (S.nstrips) {
    rc = upload_strip_tables(P, S);
    if (rc) return rc;
    HIPCHK(P, hipMemset(S.swet, 0, (size_t)S.nstrips));
  }
  if (S.hzA) {
    const int blocks = (int)std::min<size_t>((N + 255) / 256, 65536);
    sw_hz_init<<<blocks, 256, 0, P->stream>>>(S.hzA, S.hzB, N);
    HIPCHK(P, hipGetLastError());
    HIPCHK(P, hipStreamSynchronize(P->stream));
  }
  // zero state (A and B, contiguous) and max depth; the static layers sit between them
  HIPCHK(P, hipMemset(S.h, 0, (size_t)((char *)S.z - (char *)S.h)));
  HIPCHK(P, hipMemset(S.hmax, 0, N * 4));
  HIPCHK(P, hipDeviceSynchronize());
  std::vector<Inlet> hi(std::max(nInlets, 1));
  for (int m = 0; m < nInlets; ++m) {
    const int k = inlets[m], i = k % nx, j = k / nx;
    hi[m].cell = j * S.pitch + i;
    hi[m].q = 0.f;
    hi[m].h = 0.f;
    hi[m].pad = 0.f;
    hi[m].vol = 0.0;
  }
  HIPCHK(P, hipMemcpy(S.inl, hi.data(), sizeof(Inlet) * hi.size(), hipMemcpyHostToDevice));
  HIPCHK(P, hipMemset(S.ispeed, 0, (size_t)std::max(nInlets, 1) * 4));
  HIPCHK(P, hipMemset(S.iapp, 0, (size_t)std::max(nInlets, 1) * 16));
  StepRec &r = S.hrec;
  r = StepRec{};
  r.rainArea = rainArea * p->dx * p->dx;
  r.rainCells = rainField ? -1 : open;
  r.openCells = open;
  r.levels = p->lts;
  r.dirty = 1;
  r.stage = -1e30f;   // no stage until setStage
  rc = upload_rec(P, slot);
  if (rc) return rc;
  HIPCHK(P, hipStreamSynchronize(P->stream));
  P->grp.slots.clear();  // force a rebuild
  *slotOut = slot;
  return FS_OK;
}

int set_state(fs_pool *P, int slot, const double *h, const double *hu, const double *hv, int bookInitial) {
  int rc = valid_slot(P, slot);
  if (rc) return rc;
  Slot &S = P->slots[slot];
  const size_t N = (size_t)S.nx * S.ny;
  std::vector<float> th(N), tu(N), tv(N);
  double vol = 0.0;
  for (size_t k = 0; k < N; ++k) {
    const bool open = S.hmask[k] & MASK_OPEN;
    const double d = h ? h[k] : 0.0;
    th[k] = open ? (float)std::max(0.0, d) : 0.f;  // setDepth: mask ? max(0, fn) : 0
    tu[k] = hu && open ? (float)hu[k] : 0.f;        // setVelocity: h * u, so 0 in walls
    tv[k] = hv && open ? (float)hv[k] : 0.f;
    if (open) vol += (double)th[k];
  }
  if (S.nstrips) {
    std::vector<uint8_t> wet(S.nstrips, 0);
    for (int j = 0; j < S.ny; ++j)
      for (int i = 0; i < S.nx; ++i) {
        const size_t k = (size_t)j * S.nx + i;
        if (th[k] != 0.f || tu[k] != 0.f || tv[k] != 0.f) wet[strip_of(S, i, j)] = 1;
      }
    HIPCHK(P, hipMemcpyAsync(S.swet, wet.data(), (size_t)S.nstrips, hipMemcpyHostToDevice, P->stream));
    HIPCHK(P, hipStreamSynchronize(P->stream));
    HIPCHK(P, hipMemsetAsync(S.sact, 1, (size_t)S.nstrips, P->stream));
  }
  HIPCHK(P, hipMemcpyAsync(S.h, th.data(), N * 4, hipMemcpyHostToDevice, P->stream));
  HIPCHK(P, hipMemcpyAsync(S.hu, tu.data(), N * 4, hipMemcpyHostToDevice, P->stream));
  HIPCHK(P, hipMemcpyAsync(S.hv, tv.data(), N * 4, hipMemcpyHostToDevice, P->stream));
  HIPCHK(P, hipMemsetAsync(S.w, 0, N * 4, P->stream));
  rc = download_recs(P, {slot});
  if (rc) return rc;
  if (bookInitial) {
    S.hrec.rain = S.hrec.infil = S.hrec.outflow = S.hrec.clipped = S.hrec.exchanged = 0.0;
    S.volInitial = vol * S.p.dx * S.p.dx;
  }
  S.hrec.dirty = 1;
  rc = upload_rec(P, slot);
  if (rc) return rc;
  HIPCHK(P, hipStreamSynchronize(P->stream));
  return FS_OK;
}

}  // namespace

// plain copy of n float4 (16 B) elements: the achievable-HBM reference and the byte-counter
// calibration for the profiles
// depth and speed per cell for rendering; grid-stride over the nx*ny interior cells
__global__ __launch_bounds__(256) void sw_pack(const float *__restrict__ h, const float *__restrict__ hu,
                                               const float *__restrict__ hv, float hdry, size_t n,
                                               float2 *__restrict__ out) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
    const float d = h[k], qu = hu[k], qv = hv[k];
    out[k] = make_float2(d, d > hdry ? sqrtf(qu * qu + qv * qv) / d : 0.f);
  }
}

__global__ __launch_bounds__(256) void fs_copy_kernel(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

// ============================================================================ C ABI
extern "C" {

const char *fs_version(void) { return "floodsim_hip 0.1 gfx950"; }

int fs_copy_probe(size_t bytes, int32_t iters, double *gbps, double *ms) {
  if (bytes < 1024 || iters < 1) return fail(nullptr, FS_E_ARG, "bad copy probe size");
  const size_t n = bytes / 16;
  float4 *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipStream_t st = nullptr;
  // every exit releases what was made (a failed call leaves the rest unmade)
  auto done = [&](int rc) {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
  };
  hipError_t e = hipSuccess;
#define PROBE(call) \
  if ((e = (call)) != hipSuccess) return done(fail(nullptr, FS_E_HIP, "%s: %s", #call, hipGetErrorString(e)))
  PROBE(hipMalloc(&a, n * 16));
  PROBE(hipMalloc(&b, n * 16));
  PROBE(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  PROBE(hipMemsetAsync(a, 1, n * 16, st));
  PROBE(hipEventCreate(&e0));
  PROBE(hipEventCreate(&e1));
  const int blocks = 256 * 8 * 4;
  fs_copy_kernel<<<blocks, 256, 0, st>>>(a, b, n);
  PROBE(hipEventRecord(e0, st));
  for (int i = 0; i < iters; ++i) fs_copy_kernel<<<blocks, 256, 0, st>>>(a, b, n);
  PROBE(hipEventRecord(e1, st));
  PROBE(hipEventSynchronize(e1));
  PROBE(hipGetLastError());
  float t = 0.f;
  PROBE(hipEventElapsedTime(&t, e0, e1));
#undef PROBE
  const double per = t / iters;
  if (ms) *ms = per;
  if (gbps) *gbps = 2.0 * (double)(n * 16) / (per * 1e-3) / 1e9;
  return done(FS_OK);
}

int fs_pool_create(size_t hbmBudgetBytes, fs_pool_t *out) {
  if (!out) return fail(nullptr, FS_E_ARG, "out is NULL");
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev < 1)
    return fail(nullptr, FS_E_HIP, "no HIP device: %s", e != hipSuccess ? hipGetErrorString(e) : "count 0");
  fs_pool *P = new fs_pool();
  P->budget = hbmBudgetBytes;
  if (hipGetDevice(&P->device) != hipSuccess) P->device = 0;
  // FS_NO_GRAPHS=1: plain launches instead of captured K-step graphs (profilers that trace
  // kernels do not follow graph launches here)
  const char *ng = getenv("FS_NO_GRAPHS");
  if (ng && ng[0] && ng[0] != '0') P->useGraphs = false;
  e = hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete P; return fail(nullptr, FS_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e)); }
  for (auto &ev : P->ev) {
    e = hipEventCreate(&ev);
    if (e != hipSuccess) {
      for (auto &made : P->ev)
        if (made) (void)hipEventDestroy(made);
      (void)hipStreamDestroy(P->stream);
      delete P;
      return fail(nullptr, FS_E_HIP, "hipEventCreate: %s", hipGetErrorString(e));
    }
  }
  *out = P;
  return FS_OK;
}

void fs_pool_destroy(fs_pool_t P) {
  if (!P) return;
  (void)hipStreamSynchronize(P->stream);
  drop_graph(P->grp);
  for (auto &S : P->slots) {
    if (S.used && S.mem) (void)hipFree(S.mem);
    if (S.used && S.bmem) (void)hipFree(S.bmem);
    if (S.used && S.smem) (void)hipFree(S.smem);
    if (S.used && S.rmem) (void)hipFree(S.rmem);
  }
  if (P->grp.dinst) (void)hipFree(P->grp.dinst);
  if (P->grp.dwork) (void)hipFree(P->grp.dwork);
  if (P->grp.p1) (void)hipFree(P->grp.p1);
  if (P->grp.p2) (void)hipFree(P->grp.p2);
  if (P->drec) (void)hipFree(P->drec);
  if (P->dtarget) (void)hipFree(P->dtarget);
  for (auto &ev : P->ev)
    if (ev) (void)hipEventDestroy(ev);
  for (auto &ev : P->tev) (void)hipEventDestroy(ev);
  for (auto &ev : P->spanEv) (void)hipEventDestroy(ev);
  if (P->dhist) (void)hipFree(P->dhist);
  if (P->pack) (void)hipFree(P->pack);
  if (P->stream) (void)hipStreamDestroy(P->stream);
  delete P;
}

const char *fs_pool_last_error(fs_pool_t P) { return P ? P->err.c_str() : g_create_error.c_str(); }

size_t fs_pool_bytes_needed(const fs_params *p, int32_t withInfil, int32_t withRainField) {
  return p ? slot_bytes(p, withInfil != 0, withRainField != 0, 0) : 0;
}

size_t fs_pool_bytes_used(fs_pool_t P) { return P ? P->used : 0; }

int fs_pool_add(fs_pool_t P, const fs_params *p, const double *z, const double *manning, double manningScalar,
                const uint8_t *mask, const double *infil, const double *rainField, const int32_t *inlets,
                int32_t nInlets, int32_t *slot) {
  if (!P || !slot) return fail(P, FS_E_ARG, "pool or slot is NULL");
  return add_slot(P, p, z, manning, manningScalar, mask, infil, rainField, inlets, nInlets, slot);
}

int fs_pool_set_rows(fs_pool_t P, int32_t slot, int32_t rows) {
  int rc = valid_slot(P, slot);
  if (rc) return rc;
  Slot &S = P->slots[slot];
  if (S.L) return fail(P, FS_E_STATE, "rows of a local-time-stepping instance are fixed at creation");
  if (rows < kMinRows || rows > 4096) return fail(P, FS_E_ARG, "rows must be in [%d, 4096] (got %d)", kMinRows, rows);
  if (rows == S.rows) return FS_OK;
  HIPCHK(P, hipStreamSynchronize(P->stream));
  S.p.rows = rows;
  S.rows = strip_rows(&S.p);
  S.nstrips = units_of(&S.p);
  rc = upload_strip_tables(P, S);
  if (rc) return rc;
  // every strip steps once: the next second stage sets the quiet flags of the new strips
  HIPCHK(P, hipMemset(S.swet, 1, (size_t)S.nstrips));
  HIPCHK(P, hipDeviceSynchronize());
  P->grp.slots.clear();  // the work list changes: rebuild the group (and its graphs)
  return FS_OK;
}

int fs_pool_remove(fs_pool_t P, int32_t slot) {
  int rc = valid_slot(P, slot);
  if (rc) return rc;
  Slot &S = P->slots[slot];
  HIPCHK(P, hipStreamSynchronize(P->stream));
  HIPCHK(P, hipFree(S.mem));
  if (S.bmem) HIPCHK(P, hipFree(S.bmem));
  if (S.smem) HIPCHK(P, hipFree(S.smem));
  if (S.rmem) HIPCHK(P, hipFree(S.rmem));
  P->used -= S.bytes + S.bbytes + S.sbytes + S.rbytes;
  S = Slot();
  P->grp.slots.clear();
  drop_graph(P->grp);
  return FS_OK;
}

int fs_pool_set_state(fs_pool_t P, int32_t slot, const double *h, const double *hu, const double *hv,
                      int32_t bookInitial) {
  return set_state(P, slot, h, hu, hv, bookInitial);
}

int fs_pool_set_rain(fs_pool_t P, int32_t slot, double rate) {
  int rc = valid_slot(P, slot);
  if (rc) return rc;
  Slot &S = P->slots[slot];
  const float r = (float)rate;
  HIPCHK(P, hipMemcpyAsync(&P->drec[slot].rainRate, &r, sizeof r, hipMemcpyHostToDevice, P->stream));
  HIPCHK(P, hipStreamSynchronize(P->stream));
  S.hrec.rainRate = r;
  return FS_OK;
}

// water-level boundary: the stage cells (interior edge cells whose outward ghosts hold the stage;
// solver-gpu.js:125-133) and the stage (setStage, solver-gpu.js:554; stored relative to the bed minimum
// as solver-gpu.js:296 does)
int fs_pool_set_stage_cells(fs_pool_t P, int32_t slot, const int32_t *cells, int32_t n) {
  int rc = valid_slot(P, slot);
  if (rc) return rc;
  Slot &S = P->slots[slot];
  if (n < 0 || (n && !cells)) return fail(P, FS_E_ARG, "stage cells: bad list");
  const int64_t N = (int64_t)S.nx * S.ny;
  for (int q = 0; q < n; ++q)
    if (cells[q] < 0 || cells[q] >= N) return fail(P, FS_E_ARG, "stage cell %d outside the grid", cells[q]);
  for (auto &b : S.hmask) b &= (uint8_t)~MASK_STAGE;
  for (int q = 0; q < n; ++q) S.hmask[cells[q]] |= MASK_STAGE;
  S.stageCells = n;
  P->grp.slots.clear();   // the instance's kernel variant changes
  drop_graph(P->grp);
  HIPCHK(P, hipStreamSynchronize(P->stream));
  HIPCHK(P, hipMemcpy(S.mask, S.hmask.data(), (size_t)N, hipMemcpyHostToDevice));
  return upload_strip_rain(P, S);
}

int fs_pool_set_stage(fs_pool_t P, int32_t slot, double z) {
  int rc = valid_slot(P, slot);
  if (rc) return rc;
  Slot &S = P->slots[slot];
  const float st = (float)(z - S.zref);
  HIPCHK(P, hipMemcpyAsync(&P->drec[slot].stage, &st, sizeof st, hipMemcpyHostToDevice, P->stream));
  HIPCHK(P, hipStreamSynchronize(P->stream));
  S.hrec.stage = st;
  return FS_OK;
}

// flood barriers (o.crest, solver-gpu.js:18-20, 144-151): faces (i, j, axis) between two interior cells,
// crest z (absolute; non-finite = not deployed), weir coefficient C, ground under the barrier (absolute;
// non-finite = unknown, no load). A new list replaces the old one and starts its books afresh.
int fs_pool_set_crests(fs_pool_t P, int32_t slot, const int32_t *fi, const int32_t *fj, const int32_t *axis,
                       const double *z, const double *C, const double *ground, int32_t n) {
  int rc = valid_slot(P, slot);
  if (rc) return rc;
  Slot &S = P->slots[slot];
  if (n < 0 || (n && (!fi || !fj || !axis || !z || !C || !ground))) return fail(P, FS_E_ARG, "barriers: bad lists");
  if (n && (int64_t)S.nx * S.ny > kMax16BCells)
    return fail(P, FS_E_ARG, "grid too large for flood barriers: %d x %d (the 16-byte crest plane holds at most 2^28 "
                "cells)", S.nx, S.ny);
  for (int m = 0; m < n; ++m) {
    const bool ax = axis[m] != 0;
    if (fi[m] < (ax ? 0 : 1) || fi[m] >= S.nx || fj[m] < (ax ? 1 : 0) || fj[m] >= S.ny)
      return fail(P, FS_E_ARG, "barrier %d: face (%d, %d, axis %d) is not between two interior cells", m, fi[m], fj[m],
                  axis[m]);
  }
  HIPCHK(P, hipStreamSynchronize(P->stream));
  if (S.bmem) {
    HIPCHK(P, hipFree(S.bmem));
    P->used -= S.bbytes;
    S.bmem = nullptr; S.bbytes = 0; S.nbar = 0;
    S.crest = nullptr; S.bidx = nullptr; S.barFace = nullptr; S.barGround = nullptr; S.barFlux = nullptr;
    S.barBook = nullptr;
    S.hcrest.clear(); S.bars.clear();
  }
  P->grp.slots.clear();   // the instance descriptors change
  drop_graph(P->grp);
  if (!n) return FS_OK;
  const size_t N = (size_t)S.nx * S.ny;
  const size_t b = align256(N * 16) + align256(N * 8) + align256((size_t)n * 16) + 2 * align256((size_t)n * 4) +
                   align256((size_t)n * 72);
  if (P->budget && P->used + b > P->budget) return fail(P, FS_E_BUDGET, "barriers: HBM budget spent");
  HIPCHK(P, hipMalloc(&S.bmem, b));
  S.bbytes = b;
  P->used += b;
  char *q = S.bmem;
  auto carve = [&](size_t bytes) { char *r = q; q += align256(bytes); return r; };
  S.crest = (float4 *)carve(N * 16);
  S.bidx = (int2 *)carve(N * 8);
  S.barFace = (int4 *)carve((size_t)n * 16);
  S.barGround = (float *)carve((size_t)n * 4);
  S.barFlux = (float *)carve((size_t)n * 4);
  S.barBook = (double *)carve((size_t)n * 72);
  S.nbar = n;
  S.hcrest.assign(N, make_float4(-1e30f, 0.f, -1e30f, 0.f));
  std::vector<int2> bi(N, make_int2(-1, -1));
  std::vector<int4> bf(n);
  std::vector<float> gr(n);
  std::vector<double> book((size_t)n * 9, 0.0);
  S.bars.resize(n);
  for (int m = 0; m < n; ++m) {
    const int k = fj[m] * S.pitch + fi[m];
    const bool ax = axis[m] != 0;
    S.bars[m] = make_int4(fi[m], fj[m], ax ? 1 : 0, k);
    bf[m] = make_int4(k, k - (ax ? S.pitch : 1), ax ? 1 : 0, 0);
    gr[m] = std::isfinite(ground[m]) ? (float)(ground[m] - S.zref) : -1e30f;
    if (ax) { bi[k].y = m; S.hcrest[k].w = (float)C[m]; } else { bi[k].x = m; S.hcrest[k].y = (float)C[m]; }
    book[9 * (size_t)m + 3] = -1e30;   // level: never wet
    book[9 * (size_t)m + 8] = -1.0;    // first overtopping: never
  }
  HIPCHK(P, hipMemcpy(S.bidx, bi.data(), N * 8, hipMemcpyHostToDevice));
  HIPCHK(P, hipMemcpy(S.barFace, bf.data(), (size_t)n * 16, hipMemcpyHostToDevice));
  HIPCHK(P, hipMemcpy(S.barGround, gr.data(), (size_t)n * 4, hipMemcpyHostToDevice));
  HIPCHK(P, hipMemset(S.barFlux, 0, (size_t)n * 4));
  HIPCHK(P, hipMemcpy(S.barBook, book.data(), (size_t)n * 72, hipMemcpyHostToDevice));
  return fs_pool_set_crest_levels(P, slot, z);
}

// Green-Ampt soil (o.ga: psi dtheta, Fmax, pervious share, where perv > 0; Fmax capped at 1e6 m) and GI
// stores (o.gi: intake rate, capacity, drain rate, where cap > 0), per interior cell (solver-cpu.js:94-96);
// both planes live in one allocation of their own, the infiltrated depth F and the stored depth start at 0
int soil_planes(fs_pool *P, Slot &S) {
  if (S.smem) return FS_OK;
  if ((int64_t)S.nx * S.ny > kMax16BCells)
    return fail(P, FS_E_ARG, "grid too large for Green-Ampt soil or GI stores: %d x %d (their 16-byte planes hold at "
                "most 2^28 cells)", S.nx, S.ny);
  const size_t N = (size_t)S.nx * S.ny, b = 2 * align256(N * 16);
  if (P->budget && P->used + b > P->budget) return fail(P, FS_E_BUDGET, "soil/GI planes: HBM budget spent");
  HIPCHK(P, hipMalloc(&S.smem, b));
  S.sbytes = b;
  P->used += b;
  return FS_OK;
}

int fs_pool_set_ga(fs_pool_t P, int32_t slot, const double *psiD, const double *fmax, const double *perv) {
  int rc = valid_slot(P, slot);
  if (rc) return rc;
  Slot &S = P->slots[slot];
  if (!psiD || !fmax || !perv) return fail(P, FS_E_ARG, "Green-Ampt: NULL layer");
  HIPCHK(P, hipStreamSynchronize(P->stream));
  if ((rc = soil_planes(P, S))) return rc;
  const size_t N = (size_t)S.nx * S.ny;
  std::vector<float4> v(N, make_float4(0.f, 0.f, 0.f, 0.f));
  for (size_t k = 0; k < N; ++k)
    if (perv[k] > 0) v[k] = make_float4((float)psiD[k], (float)std::min(fmax[k], 1e6), (float)perv[k], 0.f);
  S.soil = (float4 *)S.smem;
  HIPCHK(P, hipMemcpy(S.soil, v.data(), N * 16, hipMemcpyHostToDevice));
  P->grp.slots.clear();   // the instance's kernel variant changes
  drop_graph(P->grp);
  return FS_OK;
}

int fs_pool_set_gi(fs_pool_t P, int32_t slot, const double *rate, const double *cap, const double *drain) {
  int rc = valid_slot(P, slot);
  if (rc) return rc;
  Slot &S = P->slots[slot];
  if (!rate || !cap || !drain) return fail(P, FS_E_ARG, "GI: NULL layer");
  HIPCHK(P, hipStreamSynchronize(P->stream));
  if ((rc = soil_planes(P, S))) return rc;
  const size_t N = (size_t)S.nx * S.ny;
  std::vector<float4> v(N, make_float4(0.f, 0.f, 0.f, 0.f));
  for (size_t k = 0; k < N; ++k)
    if (cap[k] > 0) v[k] = make_float4((float)rate[k], (float)cap[k], (float)drain[k], 0.f);
  S.gi = (float4 *)(S.smem + align256(N * 16));
  HIPCHK(P, hipMemcpy(S.gi, v.data(), N * 16, hipMemcpyHostToDevice));
  P->grp.slots.clear();
  drop_graph(P->grp);
  return FS_OK;
}

// setCrestLevels (solver-gpu.js:486-488): new crest per barrier face, in the order given
int fs_pool_set_crest_levels(fs_pool_t P, int32_t slot, const double *z) {
  int rc = valid_slot(P, slot);
  if (rc) return rc;
  Slot &S = P->slots[slot];
  if (!S.nbar) return FS_OK;
  if (!z) return fail(P, FS_E_ARG, "crest levels: NULL");
  for (int m = 0; m < S.nbar; ++m) {
    const float c = std::isfinite(z[m]) ? (float)(z[m] - S.zref) : -1e30f;
    float4 &cr = S.hcrest[S.bars[m].w];
    if (S.bars[m].z) cr.z = c; else cr.x = c;
  }
  HIPCHK(P, hipStreamSynchronize(P->stream));
  HIPCHK(P, hipMemcpy(S.crest, S.hcrest.data(), S.hcrest.size() * 16, hipMemcpyHostToDevice));
  return FS_OK;
}

// readBarriers (solver-gpu.js:490-504): 9 per face: plus, minus (m^3), peak (m^2/s), level (m, absolute;
// -inf = never wet), load (kN/m), tPeak, tLevel, tLoad, tOver (s; NaN = never overtopped)
int fs_pool_read_barriers(fs_pool_t P, int32_t slot, double *out) {
  int rc = valid_slot(P, slot);
  if (rc) return rc;
  Slot &S = P->slots[slot];
  if (!S.nbar) return FS_OK;
  if (!out) return fail(P, FS_E_ARG, "read barriers: NULL");
  HIPCHK(P, hipStreamSynchronize(P->stream));
  HIPCHK(P, hipMemcpy(out, S.barBook, (size_t)S.nbar * 72, hipMemcpyDeviceToHost));
  for (int m = 0; m < S.nbar; ++m) {
    double *b = out + 9 * (size_t)m;
    b[3] = b[3] > -1e29 ? b[3] + S.zref : -INFINITY;
    if (b[8] < 0.0) b[8] = NAN;
  }
  return FS_OK;
}

int fs_pool_set_inlet_rates(fs_pool_t P, int32_t slot, const double *q, int32_t n) {
  int rc = valid_slot(P, slot);
  if (rc) return rc;
  Slot &S = P->slots[slot];
  if (n != S.nin || (n && !q)) return fail(P, FS_E_ARG, "expected %d inlet rates, got %d", S.nin, n);
  if (!n) return FS_OK;
  std::vector<Inlet> hi(n);
  HIPCHK(P, hipMemcpyAsync(hi.data(), S.inl, sizeof(Inlet) * n, hipMemcpyDeviceToHost, P->stream));
  HIPCHK(P, hipStreamSynchronize(P->stream));
  for (int m = 0; m < n; ++m) hi[m].q = (float)q[m];
  HIPCHK(P, hipMemcpyAsync(S.inl, hi.data(), sizeof(Inlet) * n, hipMemcpyHostToDevice, P->stream));
  HIPCHK(P, hipStreamSynchronize(P->stream));
  return FS_OK;
}

static int read_inlets(fs_pool *P, int slot, double *vol, double *h, int n) {
  int rc = valid_slot(P, slot);
  if (rc) return rc;
  Slot &S = P->slots[slot];
  if (n != S.nin) return fail(P, FS_E_ARG, "expected %d inlets, got %d", S.nin, n);
  if (!n) return FS_OK;
  std::vector<Inlet> hi(n);
  HIPCHK(P, hipMemcpyAsync(hi.data(), S.inl, sizeof(Inlet) * n, hipMemcpyDeviceToHost, P->stream));
  HIPCHK(P, hipStreamSynchronize(P->stream));
  // one copy of the inlet records: the volumes since the last read and the depths the last exchange left;
  // both restart at zero, as the reference clears its INX record after the copy (solver-gpu.js:566-569)
  for (int m = 0; m < n; ++m) {
    if (vol) vol[m] = hi[m].vol;
    if (h) h[m] = hi[m].h;
    S.exchangedRead += hi[m].vol;
    hi[m].vol = 0.0;
    hi[m].h = 0.f;
  }
  HIPCHK(P, hipMemcpyAsync(S.inl, hi.data(), sizeof(Inlet) * n, hipMemcpyHostToDevice, P->stream));
  HIPCHK(P, hipStreamSynchronize(P->stream));
  return FS_OK;
}

int fs_read_inlet_approach(fs_handle s, float *app, int32_t n) {
  if (!s || (n > 0 && !app)) return FS_E_ARG;
  fs_pool *P = s->pool;
  Slot &S = P->slots[s->slot];
  if (n != S.nin) return fail(P, FS_E_ARG, "expected %d inlets, got %d", S.nin, n);
  if (!n) return FS_OK;
  HIPCHK(P, hipStreamSynchronize(P->stream));
  HIPCHK(P, hipMemcpy(app, S.iapp, (size_t)n * 16, hipMemcpyDeviceToHost));
  return FS_OK;
}

int fs_pool_advance(fs_pool_t P, const int32_t *slots, const double *tTargets, int32_t n, int32_t K,
                    fs_clock *clks) {
  if (!P || (n > 0 && (!slots || !tTargets))) return fail(P, FS_E_ARG, "bad arguments");
  int rc = advance_group(P, slots, tTargets, n, K, false);
  if (rc) return rc;
  if (clks)
    for (int i = 0; i < n; ++i) fill_clock(P->slots[slots[i]], &clks[i]);
  return FS_OK;
}

int fs_pool_step(fs_pool_t P, const int32_t *slots, const double *tTargets, int32_t n, int32_t K) {
  if (!P || (n > 0 && (!slots || !tTargets))) return fail(P, FS_E_ARG, "bad arguments");
  return advance_group(P, slots, tTargets, n, K, true);
}

int fs_pool_clock(fs_pool_t P, int32_t slot, fs_clock *clk) {
  int rc = valid_slot(P, slot);
  if (rc) return rc;
  rc = download_recs(P, {slot});
  if (rc) return rc;
  if (clk) fill_clock(P->slots[slot], clk);
  return FS_OK;
}

int fs_pool_read_state(fs_pool_t P, int32_t slot, double *h, double *hu, double *hv, double *hmax, fs_mass *mass) {
  int rc = valid_slot(P, slot);
  if (rc) return rc;
  Slot &S = P->slots[slot];
  const size_t N = (size_t)S.nx * S.ny;
  std::vector<float> th(N), tw(N), tmp(N);
  HIPCHK(P, hipMemcpyAsync(th.data(), S.h, N * 4, hipMemcpyDeviceToHost, P->stream));
  HIPCHK(P, hipMemcpyAsync(tw.data(), S.w, N * 4, hipMemcpyDeviceToHost, P->stream));
  HIPCHK(P, hipStreamSynchronize(P->stream));
  if (h) for (size_t k = 0; k < N; ++k) h[k] = th[k];
  auto fetch = [&](const float *src, double *dst) -> int {
    if (!dst) return FS_OK;
    HIPCHK(P, hipMemcpy(tmp.data(), src, N * 4, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < N; ++k) dst[k] = tmp[k];
    return FS_OK;
  };
  if ((rc = fetch(S.hu, hu))) return rc;
  if ((rc = fetch(S.hv, hv))) return rc;
  if ((rc = fetch(S.hmax, hmax))) return rc;
  rc = download_recs(P, {slot});
  if (rc) return rc;
  if (mass) {
    double vol = 0.0;
    for (size_t k = 0; k < N; ++k)
      if (S.hmask[k] & MASK_OPEN) vol += (double)th[k] - (double)tw[k];  // depth less its residual
    const double A = S.p.dx * S.p.dx;
    double pending = 0.0;
    if (S.nin) {
      std::vector<Inlet> hi(S.nin);
      HIPCHK(P, hipMemcpy(hi.data(), S.inl, sizeof(Inlet) * S.nin, hipMemcpyDeviceToHost));
      for (auto &q : hi) pending += q.vol;
    }
    const StepRec &r = S.hrec;
    fs_mass &m = *mass;
    m.volume = vol * A;
    m.rain = r.rain;
    m.infil = r.infil;
    m.outflow = r.outflow;
    m.exchanged = S.exchangedRead + pending;
    m.clipped = r.clipped;
    m.initial = S.volInitial;
    m.giStored = 0.0;   // water the GI stores hold (m^3; solver-gpu.js:683-689 mass.giStored)
    if (S.gi) {
      std::vector<float4> g(N);
      HIPCHK(P, hipMemcpy(g.data(), S.gi, N * 16, hipMemcpyDeviceToHost));
      for (size_t k = 0; k < N; ++k) m.giStored += (double)g[k].w;
      m.giStored *= A;
    }
    m.expected = S.volInitial + r.rain - r.infil - r.outflow - m.exchanged;
    // scale as readState (solver-gpu.js:685)
    const double scale = std::max(1e-30, std::max(std::fabs(m.expected),
                                                   S.volInitial + r.rain + std::max(0.0, -m.exchanged) +
                                                       std::max(0.0, -r.outflow)));
    m.rel = (m.volume - m.expected) / scale;
    m.relExClip = (m.volume - m.expected - m.clipped) / scale;
  }
  return FS_OK;
}

int fs_pool_dominant_kernel(fs_pool_t P, double *ms_total, double *cells_total, int64_t *launches) {
  if (!P) return FS_E_ARG;
  int rc = collect_timing(P);
  if (rc) return rc;
  if (ms_total) *ms_total = P->domMs;
  if (cells_total) *cells_total = P->domCells;
  if (launches) *launches = P->domLaunches;
  P->domMs = P->domCells = 0.0;
  P->domLaunches = 0;
  return FS_OK;
}

int fs_pool_debug(fs_pool_t P, int32_t what, int32_t value) {
  if (!P) return FS_E_ARG;
  if (what != FS_DEBUG_NO_STEPS) return fail(P, FS_E_ARG, "fs_pool_debug: unknown switch %d", what);
  HIPCHK(P, hipStreamSynchronize(P->stream));
  P->noSteps = value != 0;
  return FS_OK;
}

int fs_pool_set_timing(fs_pool_t P, int32_t on) {
  if (!P) return FS_E_ARG;
  HIPCHK(P, hipStreamSynchronize(P->stream));
  P->nTev = 0;
  P->spans.clear();
  P->nSpanEv = 0;
  P->nCycles = 0;
  P->domMs = P->domCells = 0.0;
  P->domLaunches = 0;
  P->timing = on != 0;
  P->ms1 = P->ms2 = 0.0;
  P->nTimed = 0;
  return FS_OK;
}

int fs_pool_kernel_ms(fs_pool_t P, double *s1, double *s2, int64_t *launches) {
  if (!P) return FS_E_ARG;
  int rc = collect_timing(P);
  if (rc) return rc;
  if (s1) *s1 = P->nTimed ? P->ms1 / P->nTimed : -1.0;
  if (s2) *s2 = P->nTimed ? P->ms2 / P->nTimed : -1.0;
  if (launches) *launches = P->nTimed;
  P->ms1 = P->ms2 = 0.0;
  P->nTimed = 0;
  return FS_OK;
}

// ---------------------------------------------------------------- single solver
int fs_create(const fs_params *p, const double *z, const double *manning, double manningScalar, const uint8_t *mask,
              const double *infil, const double *rainField, const int32_t *inlets, int32_t nInlets, fs_handle *out) {
  if (!out) return fail(nullptr, FS_E_ARG, "out is NULL");
  *out = nullptr;
  fs_pool *P = nullptr;
  int rc = fs_pool_create(0, &P);
  if (rc) return rc;
  int32_t slot = -1;
  rc = add_slot(P, p, z, manning, manningScalar, mask, infil, rainField, inlets, nInlets, &slot);
  if (rc) {
    g_create_error = P->err;
    fs_pool_destroy(P);
    return rc;
  }
  fs_solver *s = new fs_solver();
  s->pool = P;
  s->slot = slot;
  *out = s;
  return FS_OK;
}

void fs_destroy(fs_handle s) {
  if (!s) return;
  fs_pool_destroy(s->pool);
  delete s;
}

const char *fs_last_error(fs_handle s) { return s ? s->pool->err.c_str() : g_create_error.c_str(); }

fs_pool_t fs_solver_pool(fs_handle s) { return s ? s->pool : nullptr; }

int fs_set_state(fs_handle s, const double *h, const double *hu, const double *hv, int32_t bookInitial) {
  if (!s) return FS_E_ARG;
  return set_state(s->pool, s->slot, h, hu, hv, bookInitial);
}

int fs_set_rain(fs_handle s, double rate) { return s ? fs_pool_set_rain(s->pool, s->slot, rate) : FS_E_ARG; }
int fs_set_stage_cells(fs_handle s, const int32_t *cells, int32_t n) {
  return s ? fs_pool_set_stage_cells(s->pool, s->slot, cells, n) : FS_E_ARG;
}
int fs_set_stage(fs_handle s, double z) { return s ? fs_pool_set_stage(s->pool, s->slot, z) : FS_E_ARG; }
int fs_set_crests(fs_handle s, const int32_t *i, const int32_t *j, const int32_t *axis, const double *z,
                  const double *C, const double *ground, int32_t n) {
  return s ? fs_pool_set_crests(s->pool, s->slot, i, j, axis, z, C, ground, n) : FS_E_ARG;
}
int fs_set_crest_levels(fs_handle s, const double *z) {
  return s ? fs_pool_set_crest_levels(s->pool, s->slot, z) : FS_E_ARG;
}
int fs_read_barriers(fs_handle s, double *out) { return s ? fs_pool_read_barriers(s->pool, s->slot, out) : FS_E_ARG; }
int fs_set_ga(fs_handle s, const double *psiD, const double *fmax, const double *perv) {
  return s ? fs_pool_set_ga(s->pool, s->slot, psiD, fmax, perv) : FS_E_ARG;
}
int fs_set_gi(fs_handle s, const double *rate, const double *cap, const double *drain) {
  return s ? fs_pool_set_gi(s->pool, s->slot, rate, cap, drain) : FS_E_ARG;
}

int fs_set_rain_field(fs_handle s, const double *field) {
  if (!s || !field) return FS_E_ARG;
  fs_pool *P = s->pool;
  Slot &S = P->slots[s->slot];
  const size_t N = (size_t)S.nx * S.ny;
  HIPCHK(P, hipStreamSynchronize(P->stream));
  if (!S.rf) {
    // created without a field (the reference's multiplier is 1 everywhere until set): a plane of its
    // own, and the instance's descriptor changes (rain-field flag, the rain books by area)
    const size_t b = align256(N * 4);
    if (P->budget && P->used + b > P->budget) return fail(P, FS_E_BUDGET, "rain field plane: HBM budget spent");
    HIPCHK(P, hipMalloc(&S.rmem, b));
    S.rbytes = b;
    P->used += b;
    S.rf = (float *)S.rmem;
    S.hrf.assign(N, 1.f);
    P->grp.slots.clear();
    drop_graph(P->grp);
  }
  double area = 0.0;
  for (size_t k = 0; k < N; ++k) { S.hrf[k] = (float)field[k]; if (S.hmask[k] & MASK_OPEN) area += S.hrf[k]; }
  HIPCHK(P, hipMemcpy(S.rf, S.hrf.data(), N * 4, hipMemcpyHostToDevice));
  int rc = upload_strip_rain(P, S);
  if (rc) return rc;
  if (S.L) {   // the per-tile sums behind the LTS per-level rain books
    rc = upload_tile_tables(P, S);
    if (rc) return rc;
  }
  rc = download_recs(P, {s->slot});
  if (rc) return rc;
  S.hrec.rainArea = area * S.p.dx * S.p.dx;
  S.hrec.rainCells = -1;
  rc = upload_rec(P, s->slot);
  if (rc) return rc;
  HIPCHK(P, hipStreamSynchronize(P->stream));
  return FS_OK;
}

int fs_set_inlet_rates(fs_handle s, const double *q, int32_t n) {
  return s ? fs_pool_set_inlet_rates(s->pool, s->slot, q, n) : FS_E_ARG;
}

int fs_read_inlets(fs_handle s, double *vol, double *h, int32_t n) {
  return s ? read_inlets(s->pool, s->slot, vol, h, n) : FS_E_ARG;
}

int fs_advance(fs_handle s, double tTarget, int32_t K, fs_clock *clk) {
  if (!s) return FS_E_ARG;
  int32_t sl = s->slot;
  return fs_pool_advance(s->pool, &sl, &tTarget, 1, K, clk);
}

int fs_advance_noalloc(fs_handle s, double tTarget, int32_t K, fs_clock *clk) {
  if (!s) return FS_E_ARG;
  return advance_noalloc(s->pool, s->slot, tTarget, K, clk);
}

int fs_frame(fs_handle s, double tEnd, int32_t K, float *out, fs_clock *clk) {
  if (!s) return FS_E_ARG;
  fs_pool *P = s->pool;
  int32_t sl = s->slot;
  // K = 0: the readout alone (the reference encodes no steps then, solver-gpu.js:611)
  int rc = K > 0 ? advance_group(P, &sl, &tEnd, 1, K, true) : FS_OK;
  if (rc) return rc;
  Slot &S = P->slots[sl];
  if (out) {
    // pack (kernels.wgsl.js:1346-1355) on the device: depth and speed per interior cell, 8 B read back
    const size_t N = (size_t)S.nx * S.ny;
    if (N > P->capPack) {
      if (P->pack) HIPCHK(P, hipFree(P->pack));
      P->pack = nullptr;
      HIPCHK(P, hipMalloc(&P->pack, N * sizeof(float2)));
      P->capPack = N;
    }
    const int blocks = (int)std::min<size_t>((N + 255) / 256, 65536);
    sw_pack<<<blocks, 256, 0, P->stream>>>(S.h, S.hu, S.hv, (float)S.p.hdry, N, P->pack);
    HIPCHK(P, hipGetLastError());
    HIPCHK(P, hipMemcpyAsync(out, P->pack, N * sizeof(float2), hipMemcpyDeviceToHost, P->stream));
    HIPCHK(P, hipStreamSynchronize(P->stream));
  }
  if (clk) fill_clock(S, clk);
  return FS_OK;
}

// ---- device-state snapshot / restore (checkpoint.js:94-113 keeps the solver's buffers and its
// host-side clock and books): the slot's whole device allocation, its step record and host books
namespace {
struct SnapHeader {
  char magic[8];
  uint32_t version, headerBytes;
  uint64_t slotBytes;
  int32_t nx, ny, lts, rows, secondOrder, limiter, bc[4];
  double dx, g, cfl, hdry, hfront, dtDry;
  StepRec rec;
  double volInitial, exchangedRead, dtSum;
  int64_t dtSteps;
  // the allocations of their own that follow the main one (barriers, soil/GI planes, a rain field given
  // after creation), and the options they belong to
  uint64_t bBytes, sBytes, rBytes;
  int32_t nbar, stageCells, hasSoil, hasGi, hasRf, hazard;
};
constexpr char kSnapMagic[8] = {'F', 'S', 'S', 'N', 'A', 'P', '0', '3'};
constexpr uint32_t kSnapVersion = 3;

bool same_model(const SnapHeader &h, const Slot &S) {
  const fs_params &p = S.p;
  return h.nx == p.nx && h.ny == p.ny && h.lts == p.lts && h.secondOrder == p.secondOrder && h.limiter == p.limiter &&
         h.bc[0] == p.bc[0] && h.bc[1] == p.bc[1] && h.bc[2] == p.bc[2] && h.bc[3] == p.bc[3] && h.dx == p.dx &&
         h.g == p.g && h.cfl == p.cfl && h.hdry == p.hdry && h.hfront == p.hfront && h.dtDry == p.dtDry &&
         h.slotBytes == S.bytes && h.bBytes == S.bbytes && h.sBytes == S.sbytes && h.rBytes == S.rbytes &&
         h.nbar == S.nbar && h.stageCells == S.stageCells && h.hasSoil == (S.soil != nullptr) &&
         h.hasGi == (S.gi != nullptr) && h.hasRf == (S.rf != nullptr) && h.hazard == (S.hzA != nullptr);
}

size_t snap_bytes(const Slot &S) { return sizeof(SnapHeader) + S.bytes + S.bbytes + S.sbytes + S.rbytes; }
}  // namespace

size_t fs_pool_snapshot_bytes(fs_pool_t P, int32_t slot) {
  if (valid_slot(P, slot)) return 0;
  return snap_bytes(P->slots[slot]);
}

int fs_pool_snapshot(fs_pool_t P, int32_t slot, void *out, size_t bytes) {
  int rc = valid_slot(P, slot);
  if (rc) return rc;
  Slot &S = P->slots[slot];
  if (!out || bytes < snap_bytes(S)) return fail(P, FS_E_ARG, "snapshot needs %zu bytes", snap_bytes(S));
  rc = download_recs(P, {slot});
  if (rc) return rc;
  SnapHeader h{};
  std::memcpy(h.magic, kSnapMagic, 8);
  h.version = kSnapVersion;
  h.headerBytes = sizeof(SnapHeader);
  h.slotBytes = S.bytes;
  h.nx = S.p.nx; h.ny = S.p.ny; h.lts = S.p.lts; h.rows = S.p.rows; h.secondOrder = S.p.secondOrder;
  h.limiter = S.p.limiter;
  for (int k = 0; k < 4; ++k) h.bc[k] = S.p.bc[k];
  h.dx = S.p.dx; h.g = S.p.g; h.cfl = S.p.cfl; h.hdry = S.p.hdry; h.hfront = S.p.hfront; h.dtDry = S.p.dtDry;
  h.rec = S.hrec;
  h.volInitial = S.volInitial; h.exchangedRead = S.exchangedRead; h.dtSum = S.dtSum; h.dtSteps = S.dtSteps;
  h.bBytes = S.bbytes; h.sBytes = S.sbytes; h.rBytes = S.rbytes;
  h.nbar = S.nbar; h.stageCells = S.stageCells; h.hasSoil = S.soil != nullptr; h.hasGi = S.gi != nullptr;
  h.hasRf = S.rf != nullptr; h.hazard = S.hzA != nullptr;
  std::memcpy(out, &h, sizeof h);
  char *q = (char *)out + sizeof h;
  HIPCHK(P, hipMemcpy(q, S.mem, S.bytes, hipMemcpyDeviceToHost));
  q += S.bytes;
  if (S.bbytes) HIPCHK(P, hipMemcpy(q, S.bmem, S.bbytes, hipMemcpyDeviceToHost));
  q += S.bbytes;
  if (S.sbytes) HIPCHK(P, hipMemcpy(q, S.smem, S.sbytes, hipMemcpyDeviceToHost));
  q += S.sbytes;
  if (S.rbytes) HIPCHK(P, hipMemcpy(q, S.rmem, S.rbytes, hipMemcpyDeviceToHost));
  return FS_OK;
}

int fs_pool_restore(fs_pool_t P, int32_t slot, const void *in, size_t bytes) {
  int rc = valid_slot(P, slot);
  if (rc) return rc;
  Slot &S = P->slots[slot];
  SnapHeader h;
  if (!in || bytes < sizeof h) return fail(P, FS_E_ARG, "not a snapshot (too short)");
  std::memcpy(&h, in, sizeof h);
  if (std::memcmp(h.magic, kSnapMagic, 8) != 0 || h.version != kSnapVersion || h.headerBytes != sizeof h)
    return fail(P, FS_E_ARG, "not a snapshot of this library version");
  if (!same_model(h, S)) return fail(P, FS_E_ARG, "snapshot of a different model (grid, options or layout)");
  if (bytes < snap_bytes(S)) return fail(P, FS_E_ARG, "snapshot truncated");
  HIPCHK(P, hipStreamSynchronize(P->stream));
  const char *q = (const char *)in + sizeof h;
  HIPCHK(P, hipMemcpy(S.mem, q, S.bytes, hipMemcpyHostToDevice));
  q += S.bytes;
  if (S.bbytes) HIPCHK(P, hipMemcpy(S.bmem, q, S.bbytes, hipMemcpyHostToDevice));
  q += S.bbytes;
  if (S.sbytes) HIPCHK(P, hipMemcpy(S.smem, q, S.sbytes, hipMemcpyHostToDevice));
  q += S.sbytes;
  if (S.rbytes) HIPCHK(P, hipMemcpy(S.rmem, q, S.rbytes, hipMemcpyHostToDevice));
  // the host mirrors of restored planes (mask bits, rain multiplier, crests) follow the device
  const size_t N = (size_t)S.nx * S.ny;
  HIPCHK(P, hipMemcpy(S.hmask.data(), S.mask, N, hipMemcpyDeviceToHost));
  if (S.rf) {
    S.hrf.resize(N);
    HIPCHK(P, hipMemcpy(S.hrf.data(), S.rf, N * 4, hipMemcpyDeviceToHost));
  }
  if (S.crest && !S.hcrest.empty()) HIPCHK(P, hipMemcpy(S.hcrest.data(), S.crest, S.hcrest.size() * 16, hipMemcpyDeviceToHost));
  // the strip tables in the restored memory belong to the snapshot's strip height
  S.p.rows = h.rows;
  S.rows = strip_rows(&S.p);
  if (S.nstrips) S.nstrips = units_of(&S.p);
  S.hrec = h.rec;
  S.volInitial = h.volInitial; S.exchangedRead = h.exchangedRead; S.dtSum = h.dtSum; S.dtSteps = h.dtSteps;
  rc = upload_rec(P, slot);
  if (rc) return rc;
  HIPCHK(P, hipStreamSynchronize(P->stream));
  P->grp.slots.clear();
  return FS_OK;
}

int fs_read_infiltration(fs_handle s, float *F) {
  if (!s || !F) return FS_E_ARG;
  fs_pool *P = s->pool;
  Slot &S = P->slots[s->slot];
  const size_t N = (size_t)S.nx * S.ny;
  if (!S.soil) {
    std::fill(F, F + N, 0.f);
    return FS_OK;
  }
  std::vector<float4> a(N);
  HIPCHK(P, hipStreamSynchronize(P->stream));
  HIPCHK(P, hipMemcpy(a.data(), S.soil, N * 16, hipMemcpyDeviceToHost));
  for (size_t k = 0; k < N; ++k) F[k] = a[k].w;   // (psi dtheta, Fmax, pervious share, F)
  return FS_OK;
}

int fs_read_hazard(fs_handle s, float *dv, float *force, float *duration, float *arrival, float *vmax, float *tWet,
                   uint8_t *arr, uint8_t *defra, uint8_t *defraUrban) {
  if (!s) return FS_E_ARG;
  fs_pool *P = s->pool;
  Slot &S = P->slots[s->slot];
  if (!S.hzA) return fail(P, FS_E_STATE, "hazard products are tracked only for a solver created with hazard = 1");
  const size_t N = (size_t)S.nx * S.ny;
  std::vector<float4> a(N), b(N);
  HIPCHK(P, hipStreamSynchronize(P->stream));
  HIPCHK(P, hipMemcpy(a.data(), S.hzA, N * 16, hipMemcpyDeviceToHost));
  HIPCHK(P, hipMemcpy(b.data(), S.hzB, N * 16, hipMemcpyDeviceToHost));
  const float nan = std::nanf("");
  for (size_t k = 0; k < N; ++k) {
    if (dv) dv[k] = a[k].x;
    if (force) force[k] = a[k].y;
    if (duration) duration[k] = a[k].z;
    if (arrival) arrival[k] = a[k].w >= 0.f ? a[k].w : nan;
    if (vmax) vmax[k] = b[k].x;
    if (tWet) tWet[k] = b[k].y >= 0.f ? b[k].y : nan;
    const int c = (int)b[k].z;
    if (arr) arr[k] = (uint8_t)(c & 7);
    if (defra) defra[k] = (uint8_t)((c >> 3) & 7);
    if (defraUrban) defraUrban[k] = (uint8_t)(c >> 6);
  }
  return FS_OK;
}

size_t fs_snapshot_bytes(fs_handle s) { return s ? fs_pool_snapshot_bytes(s->pool, s->slot) : 0; }
int fs_snapshot(fs_handle s, void *out, size_t bytes) { return s ? fs_pool_snapshot(s->pool, s->slot, out, bytes) : FS_E_ARG; }
int fs_restore(fs_handle s, const void *in, size_t bytes) { return s ? fs_pool_restore(s->pool, s->slot, in, bytes) : FS_E_ARG; }
int fs_snapshot_layout(fs_handle s, uint64_t bytes[FS_SNAP_PARTS]) {
  if (!s || !bytes) return FS_E_ARG;
  const Slot &S = s->pool->slots[s->slot];
  bytes[FS_SNAP_RECORD] = sizeof(SnapHeader);
  bytes[FS_SNAP_PLANES] = S.bytes;
  bytes[FS_SNAP_BARRIERS] = S.bbytes;
  bytes[FS_SNAP_SOIL] = S.sbytes;
  bytes[FS_SNAP_RAINFIELD] = S.rbytes;
  return FS_OK;
}

int fs_read_state(fs_handle s, double *h, double *hu, double *hv, double *hmax, fs_mass *mass) {
  return s ? fs_pool_read_state(s->pool, s->slot, h, hu, hv, hmax, mass) : FS_E_ARG;
}

int fs_clock_get(fs_handle s, fs_clock *clk) { return s ? fs_pool_clock(s->pool, s->slot, clk) : FS_E_ARG; }

int fs_lts_levels(fs_handle s, uint8_t *levels, int32_t *tx, int32_t *ty) {
  if (!s) return FS_E_ARG;
  fs_pool *P = s->pool;
  Slot &S = P->slots[s->slot];
  if (tx) *tx = S.tx;
  if (ty) *ty = S.ty;
  if (!S.L) return fail(P, FS_E_STATE, "no local time stepping on this solver");
  HIPCHK(P, hipStreamSynchronize(P->stream));
  if (levels) HIPCHK(P, hipMemcpy(levels, S.lv, (size_t)S.tx * S.ty, hipMemcpyDeviceToHost));
  return FS_OK;
}

int fs_tile_stats(fs_handle s, fs_tiles *out) {
  if (!s || !out) return FS_E_ARG;
  Slot &S = s->pool->slots[s->slot];
  int rc = download_recs(s->pool, {s->slot});
  if (rc) return rc;
  *out = fs_tiles{};
  out->cellsTotal = (int64_t)S.nx * S.ny;
  if (S.L) {   // local time stepping: every 8x8 level tile steps each cycle
    out->tiled = 1;
    out->tiles = out->update = out->flux = (int64_t)S.tx * S.ty;
    out->cellsActive = S.openCells;
    return FS_OK;
  }
  // the same rule as the descriptor's strip flags (build_group): skipping is off untiled and with GI stores
  const bool skip = S.nstrips && !S.p.untiled && !S.gi;
  out->tiled = skip ? 1 : 0;
  out->tiles = S.nstrips;
  out->update = out->flux = skip ? S.hrec.activeStrips : S.nstrips;
  out->cellsActive = skip ? S.hrec.active : S.openCells;
  return FS_OK;
}

}  // extern "C"

// flip_capi.hip - the C ABI of include/splashtank.h over HIP (host code).
//
// One Sim per tank: device buffers, one HIP stream, and a ring of step slots
// (parameter block + collider records) so consecutive steps can be enqueued
// without waiting for each other.  Mirrors the WebGPU side of the reference's
// FlipSim (src/flip/sim.js:46-547).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <atomic>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "flip_adjoint.h"
#include "flip_device.h"
#include "flip_kernels.h"
#include "splashtank.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                            \
  do {                                                                                          \
    hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess) return fail(ST_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct StepSlot {
  DParams P;
  DCols C;
};
const int RING = 64;

template <class T> int dalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) return fail(ST_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return ST_OK;
}
template <class T> void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

// A captured plain substep and what it was captured for: the device pointers
// (by value in every kernel's arguments), the host flags that select kernels,
// and the sizes that set launch grids.  A replay is valid exactly when all of
// them are the same again.
struct StepGraph {
  DGrid G;
  bool binned, solid;
  int iters;
  uint32_t N, scap, pipe;
  // which whitewater kernels the spray stage launches, and over how many
  // drip samples and air cells (host flags of the step, not device state)
  bool airOn, sprOn, dripOn;
  uint32_t ndrip, airN, sprayCap;
  uint32_t extrap;  // layers of the opt-in velocity extrapolation
  hipGraphExec_t exec;
};
const size_t GRAPH_CACHE = 4;

// One pinned staging buffer per sim, lent to one read at a time; a read that
// finds it lent, or too small while lent, gets a buffer of its own.
struct PinCache {
  std::mutex m;
  char* ptr = nullptr;
  size_t bytes = 0;
  bool lent = false;
  ~PinCache() {
    if (ptr) (void)hipHostFree(ptr);
  }
};

struct Sim {
  int dev = 0;
  hipStream_t st = nullptr;
  // sizes (sim.js:47-65)
  double hd = 0, dd = 0;
  float d = 0, h = 0, ratio = 0, rest = 0;
  int32_t nx = 0, ny = 0, nz = 0;
  uint32_t tx = 0, ty = 0, tz = 0, ntiles = 0, scap = 0, scapMax = 0, maxn = 0, ncell = 0, open = 0;
  float x0 = 0, y0 = 0, z0 = 0, x1 = 0, y1 = 0, z1 = 0;
  uint32_t N = 0;
  DGrid G{};
  // step slots
  StepSlot* dSlots = nullptr;
  StepSlot* hSlots = nullptr;
  hipEvent_t ev[RING];
  bool evUsed[RING];
  uint32_t ringPos = 0;
  StepSlot* cur = nullptr;
  DParams hp{};
  // colliders
  float* dSdf = nullptr;
  size_t sdfCap = 0, sdfN = 0;
  uint64_t sdfSig = ~0ull, sdfSigPrev = ~0ull;
  std::vector<float> cf, cfPrev;
  uint32_t ncTotal = 0;
  bool haveCols = false, solidDirty = true;
  // host state of FlipSim
  double tm = 0;
  uint32_t seed = 0;
  bool velGOn = false, nbOn = false, onlyTiles = false, begun = false;
  uint32_t initStep = 0;
  int iters = 1;
  uint32_t lcap = 0;
  uint32_t sorResident = 2048;  // SOR workgroups resident at once on this device
  uint32_t coopGrid = 0;        // resident-SOR workgroups (0: not available)
  uint32_t sorSeq = 0;          // pressure solves so far
  uint32_t sorTag = 1;          // tag of the next substep's first halo exchange (0 is never used)
  uint32_t sorMode = 0;         // this step's pressure path (st_params.sor_mode, a tuned choice for 3)
  uint32_t pipeOverride = 0;    // workgroups of the per-iteration SOR launch (0: the default)
  bool tuned = false;           // st_tune_sor has chosen a path
  StepSlot* dGraphSlot = nullptr;  // the step slot every captured substep reads
  std::vector<struct StepGraph> graphs;  // captured substeps (st_step_n with use_graph)
  uint32_t tunedMode = 0, tunedPipe = 0;
  // need, binCount and binKey already hold the marks and bins of A's positions
  // (the last g2p made them and nothing has moved a particle since)
  bool binned = false;
  NbScratch nb{};
  void* batch = nullptr;          // the batch the tank belongs to (its stream is the batch's)
  // whitewater (setSprayCap, airAlloc; sim.js:333-351)
  uint32_t sprayCap = 0;
  uint32_t airN = 0, aNx = 1, aNy = 1, aNz = 1;  // air field cells (0: not made yet)
  double airH = 1.0;
  bool airOnStep = false, sprOnStep = false;
  // opt-in velocity extrapolation: layers this step, masks for exCells cells
  uint32_t extrapOnStep = 0, exCells = 0;
  uint8_t* exMask = nullptr;
  uint32_t ndrip = 0, dripCap = 0;  // drip samples set, room for them
  bool dripOnStep = false;
  // live view (st_pack_view): packed words, sample flags, their scan, the sampled whitewater
  uint32_t* viewOut = nullptr;
  uint32_t *viewFlag = nullptr, *viewOff = nullptr, *viewSums = nullptr;
  float4* viewDisp = nullptr;
  size_t viewWords = 0;
  uint32_t viewSd = 0;
  // deep water points (st_read_deep): flags, offsets, scan sums and output,
  // sized for deepCells cells; deepPrev = deep cells counted by the last read
  uint32_t *deepFlag = nullptr, *deepOff = nullptr, *deepSums = nullptr, *deepAll = nullptr;
  float4* deepOut = nullptr;
  uint32_t deepCells = 0;
  // shared with the asynchronous reads, which set it when they are waited for
  std::shared_ptr<std::atomic<uint32_t>> deepPrev = std::make_shared<std::atomic<uint32_t>>(0u);
  std::shared_ptr<PinCache> pin;  // pinned staging of the asynchronous reads
  double* normPart = nullptr;     // st_solve_norms: per-workgroup partials, then the result
  uint32_t normCap = 0;           // workgroups normPart has room for
  hipStream_t ownStream = nullptr;  // the tank's own stream while it is in a batch

  LaunchCtx ctx() const {
    LaunchCtx L;
    L.P = &cur->P;
    L.C = &cur->C;
    L.G = G;
    L.n = N;
    L.ntiles = ntiles;
    L.scap = scap;
    L.cellGrid = std::max<uint32_t>(1u, std::min<uint32_t>(scap * 2u, 4096u));
    L.tileGrid = std::max<uint32_t>(1u, std::min<uint32_t>(scap, 8192u));
    L.pipeGrid = pipeOverride ? pipeOverride : std::min<uint32_t>((scap + 7u) & ~7u, sorResident);  // whole groups of 8 (one per XCD)
    L.coopGrid = sorMode == 1u ? 0u : coopGrid;
    L.sorStream = sorMode == 2u;
    L.sorSeq = sorSeq;
    L.binned = binned;
    L.binNext = !nbOn;  // the narrow band's reseed moves particles after g2p
    L.nbOn = nbOn;
    L.initStep = initStep;
    L.smax = sprayCap;
    L.extrap = extrapOnStep;
    L.exMask = exMask;
    return L;
  }
};

int allocSlots(Sim* s, uint32_t n) {
  size_t C = size_t(n) * 512;
  int rc;
  if ((rc = dalloc(&s->G.cnt, C))) return rc;
  if ((rc = dalloc(&s->G.ctype, C))) return rc;
  if ((rc = dalloc(&s->G.pres, C))) return rc;
  if ((rc = dalloc(&s->G.pres2, C))) return rc;
  if ((rc = dalloc(&s->G.rhs, C))) return rc;
  if ((rc = dalloc(&s->G.sten, C))) return rc;
  if ((rc = dalloc(&s->G.accV, C * 3))) return rc;
  if ((rc = dalloc(&s->G.accW, C * 3))) return rc;
  if ((rc = dalloc(&s->G.p2gHalo, size_t(n) * P2G_HALO_WORDS))) return rc;
  if (s->velGOn && (rc = dalloc(&s->G.velG, C * 3))) return rc;
  if ((rc = dalloc(&s->G.slotTile, n))) return rc;
  if ((rc = dalloc(&s->G.fstack, n))) return rc;
  if ((rc = dalloc(&s->G.slotList, n))) return rc;
  if ((rc = dalloc(&s->G.nbr, size_t(n) * 27))) return rc;
  if ((rc = dalloc(&s->G.nbrL, size_t(n) * 27))) return rc;
  if ((rc = dalloc(&s->G.coopList, n))) return rc;
  if ((rc = dalloc(&s->G.sorHalo, size_t(n) * 2 * 6 * 32))) return rc;
  // tag 0 is never a half-sweep's tag, so no entry looks written before it is
  HIPCHK(hipMemset(s->G.sorHalo, 0, size_t(n) * 2 * 6 * 32 * sizeof(uint64_t)));
  return ST_OK;
}

void freeSlotArrays(DGrid& g) {
  dfree(g.cnt); dfree(g.ctype); dfree(g.pres); dfree(g.pres2); dfree(g.rhs); dfree(g.sten);
  dfree(g.accV); dfree(g.accW); dfree(g.p2gHalo); dfree(g.velG);
  dfree(g.slotTile); dfree(g.fstack); dfree(g.slotList); dfree(g.nbr); dfree(g.nbrL);
  dfree(g.coopList); dfree(g.sorHalo);
}
void freeSlots(Sim* s) { freeSlotArrays(s->G); }

int zeroSlots(Sim* s, uint32_t from, uint32_t to) {
  if (to <= from) return ST_OK;
  size_t a = size_t(from) * 512, n = size_t(to - from) * 512;
  HIPCHK(hipMemsetAsync(s->G.cnt + a, 0, n * 4, s->st));
  HIPCHK(hipMemsetAsync(s->G.ctype + a, 0, n * 4, s->st));
  HIPCHK(hipMemsetAsync(s->G.pres + a, 0, n * 4, s->st));
  HIPCHK(hipMemsetAsync(s->G.pres2 + a, 0, n * 4, s->st));
  HIPCHK(hipMemsetAsync(s->G.rhs + a, 0, n * 4, s->st));
  HIPCHK(hipMemsetAsync(s->G.sten + a, 0, n, s->st));
  HIPCHK(hipMemsetAsync(s->G.accV + a * 3, 0, n * 12, s->st));
  HIPCHK(hipMemsetAsync(s->G.accW + a * 3, 0, n * 12, s->st));
  if (s->G.velG) HIPCHK(hipMemsetAsync(s->G.velG + a * 3, 0, n * 12, s->st));
  return ST_OK;
}

void freeNb(NbScratch& S) {
  dfree(S.deadFlag); dfree(S.deadStart); dfree(S.deadList); dfree(S.need); dfree(S.needStart); dfree(S.ent); dfree(S.entIdx);
  dfree(S.sums);
  S.ncap = S.nmax = 0;
}
// reseed scratch sized for the pool and the particle capacity
int ensureNb(Sim* s) {
  uint32_t ncap = s->scap * 512u;
  if (s->nb.deadFlag && s->nb.ncap == ncap && s->nb.nmax == s->maxn) return ST_OK;
  freeNb(s->nb);
  size_t M = s->maxn, big = std::max<size_t>(M, ncap);
  int rc;
  if ((rc = dalloc(&s->nb.deadFlag, M)) || (rc = dalloc(&s->nb.deadStart, M + 1)) || (rc = dalloc(&s->nb.deadList, M)) ||
      (rc = dalloc(&s->nb.need, ncap)) || (rc = dalloc(&s->nb.needStart, size_t(ncap) + 1)) || (rc = dalloc(&s->nb.ent, ncap)) ||
      (rc = dalloc(&s->nb.entIdx, size_t(ncap) + 1)) || (rc = dalloc(&s->nb.sums, big / 1024 + 2))) {
    freeNb(s->nb);
    return rc;
  }
  s->nb.ncap = ncap;
  s->nb.nmax = s->maxn;
  return ST_OK;
}

void destroySim(Sim* s) {
  if (!s) return;
  if (s->st) (void)hipStreamSynchronize(s->st);
  freeNb(s->nb);
  dfree(s->G.pos); dfree(s->G.vel); dfree(s->G.pid); dfree(s->G.posB); dfree(s->G.velB); dfree(s->G.pidB); dfree(s->G.binIdx);
  dfree(s->G.binKey); dfree(s->G.binRank); dfree(s->G.binCount); dfree(s->G.binStart); dfree(s->G.scanTmp);
  dfree(s->G.tmap); dfree(s->G.need); dfree(s->G.newList); dfree(s->G.wg); dfree(s->G.ctl); dfree(s->G.part); dfree(s->G.foamTab);
  dfree(s->G.mortonTiles); dfree(s->G.coopFlags); dfree(s->G.coopOffs); dfree(s->G.sorBar);
  dfree(s->G.scratch4);
  dfree(s->normPart);
  dfree(s->G.sPos); dfree(s->G.sVel); dfree(s->G.sAux); dfree(s->G.sctl); dfree(s->G.sprCnt); dfree(s->G.sprOff); dfree(s->G.sprSums);
  dfree(s->G.air); dfree(s->G.air2); dfree(s->G.airAcc);
  dfree(s->G.drip); dfree(s->G.dripFlag); dfree(s->G.dripOff); dfree(s->G.dripSums);
  dfree(s->viewOut); dfree(s->viewFlag); dfree(s->viewOff); dfree(s->viewSums); dfree(s->viewDisp);
  dfree(s->deepFlag); dfree(s->deepOff); dfree(s->deepSums); dfree(s->deepAll); dfree(s->deepOut);
  dfree(s->exMask);
  freeSlots(s);
  dfree(s->dSdf);
  dfree(s->dSlots);
  if (s->hSlots) (void)hipHostFree(s->hSlots);
  for (int k = 0; k < RING; k++)
    if (s->evUsed[k]) (void)hipEventDestroy(s->ev[k]);
  for (StepGraph& g : s->graphs) (void)hipGraphExecDestroy(g.exec);
  s->graphs.clear();
  dfree(s->dGraphSlot);
  if (s->st) (void)hipStreamDestroy(s->st);
  delete s;
}

Sim* getSim(void* p) { return (Sim*)p; }

// airAlloc (sim.js:343-351): the air field's grid, cells a few solver cells
// wide, at most about 260 thousand of them; made when first used, zeroed
int airAlloc(Sim* s) {
  const double r = std::max(2.0, std::ceil(std::cbrt(double(s->ncell) / 260000.0)));
  const uint32_t nx = uint32_t(std::ceil(double(s->nx) / r)) + 1u, ny = uint32_t(std::ceil(double(s->ny) / r)) + 1u,
                 nz = uint32_t(std::ceil(double(s->nz) / r)) + 1u, n = nx * ny * nz;
  HIPCHK(hipStreamSynchronize(s->st));
  int rc;
  if ((rc = dalloc(&s->G.air, n)) || (rc = dalloc(&s->G.air2, n)) || (rc = dalloc(&s->G.airAcc, size_t(n) * 4))) return rc;
  HIPCHK(hipMemset(s->G.air, 0, size_t(n) * 16));
  HIPCHK(hipMemset(s->G.air2, 0, size_t(n) * 16));
  HIPCHK(hipMemset(s->G.airAcc, 0, size_t(n) * 16));
  s->airH = s->hd * r;
  s->aNx = nx; s->aNy = ny; s->aNz = nz;
  s->airN = n;
  return ST_OK;
}

// A world box around colNear's region of one collider record (32 floats: a =
// sdf base and dims, b = grid origin and cell, rows of R, centre; colNear
// takes g = (R (p - c) - origin) / cell and asks lo < g < dims + hq): the
// local box [origin + lo cell, origin + (dims + hq) cell] through p = c + R^-1 l
// in f64, padded by 1e-4 of the scene's size.  The f32 test differs from the
// exact one by a few ulps of |p - c|, far inside the padding, so a point outside
// the box is one colNear turns down.  Unbounded when R is singular or a value
// is not finite.
void nearBox(const float* r, float lo, float hq, float4& wl, float4& wh) {
  const double inf = std::numeric_limits<double>::infinity();
  double M[3][3], R[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) R[i][j] = r[8 + 4 * i + j];
  const double det = R[0][0] * (R[1][1] * R[2][2] - R[1][2] * R[2][1]) - R[0][1] * (R[1][0] * R[2][2] - R[1][2] * R[2][0]) +
                     R[0][2] * (R[1][0] * R[2][1] - R[1][1] * R[2][0]);
  double l0[3], l1[3], ext = 1.0;
  bool ok = std::isfinite(det) && std::fabs(det) > 1e-12;
  for (int j = 0; j < 3; j++) {
    l0[j] = double(r[4 + j]) + double(lo) * double(r[7]);
    l1[j] = double(r[4 + j]) + (double(r[1 + j]) + double(hq)) * double(r[7]);
    ok = ok && std::isfinite(l0[j]) && std::isfinite(l1[j]) && std::isfinite(double(r[20 + j]));
    ext += std::fabs(l0[j]) + std::fabs(l1[j]) + std::fabs(double(r[20 + j]));
  }
  if (!ok) {
    wl = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
    wh = make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
    return;
  }
  // inverse by cofactors
  M[0][0] = (R[1][1] * R[2][2] - R[1][2] * R[2][1]) / det; M[0][1] = (R[0][2] * R[2][1] - R[0][1] * R[2][2]) / det;
  M[0][2] = (R[0][1] * R[1][2] - R[0][2] * R[1][1]) / det; M[1][0] = (R[1][2] * R[2][0] - R[1][0] * R[2][2]) / det;
  M[1][1] = (R[0][0] * R[2][2] - R[0][2] * R[2][0]) / det; M[1][2] = (R[0][2] * R[1][0] - R[0][0] * R[1][2]) / det;
  M[2][0] = (R[1][0] * R[2][1] - R[1][1] * R[2][0]) / det; M[2][1] = (R[0][1] * R[2][0] - R[0][0] * R[2][1]) / det;
  M[2][2] = (R[0][0] * R[1][1] - R[0][1] * R[1][0]) / det;
  const double pad = 1e-4 * ext;
  double lo3[3], hi3[3];
  for (int i = 0; i < 3; i++) {
    double mid = double(r[20 + i]), half = 0.0;
    for (int j = 0; j < 3; j++) {
      mid += M[i][j] * 0.5 * (l0[j] + l1[j]);
      half += std::fabs(M[i][j]) * 0.5 * std::fabs(l1[j] - l0[j]);
    }
    lo3[i] = mid - half - pad;
    hi3[i] = mid + half + pad;
    if (!std::isfinite(lo3[i]) || !std::isfinite(hi3[i])) { lo3[i] = -inf; hi3[i] = inf; }
  }
  wl = make_float4(float(lo3[0]), float(lo3[1]), float(lo3[2]), 0.0f);
  wh = make_float4(float(hi3[0]), float(hi3[1]), float(hi3[2]), 0.0f);
}

// the step's parameter block and collider records (sim.js:179-231) into P and C
// The largest f32 T at or below (c dt sens)^2 (1 - 1e-5), or -1 when dt or
// sens is not a positive finite number.  For a squared length s <= T the
// quotient g2p clamps, RN(RN(RN(sqrt(s)) / dt) / sens), is at most
// c sqrt(1 - 1e-5) (1 + 2^-24)^3 < c before its last rounding, hence at most c
// (c is an f32) after it; g2p's impact (c = 15) and turb (c = 0.35f, dt = 1)
// are then 0.
static float quietBelow(double c, float dt, float sens) {
  if (!(dt > 0.0f) || !(sens > 0.0f) || !std::isfinite(dt) || !std::isfinite(sens)) return -1.0f;
  const double r = c * double(dt) * double(sens);
  double t = r * r * (1.0 - 1e-5);
  if (!(t < double(FLT_MAX))) t = double(FLT_MAX);
  float f = float(t);
  if (double(f) > t) f = std::nextafter(f, 0.0f);
  return f;
}

int prepStep(Sim* s, const st_params* o, DParams& Pout, DCols& C) {
  if (s->N == 0) return fail(ST_ERR_STATE, "st_begin_step: no particles (step() does nothing when N = 0)");
  if (o->ocean_on && o->nwave > ST_OCEAN_MAX_WAVES) return fail(ST_ERR_ARG, "st_begin_step: nwave above ST_OCEAN_MAX_WAVES");
  s->nbOn = o->nb_on != 0 || o->ocean_on != 0;  // the ocean runs with the narrow band (sim.js:185-186)
  if (s->nbOn && !s->velGOn) {  // enableNB (sim.js:120-126)
    HIPCHK(hipStreamSynchronize(s->st));
    s->velGOn = true;
    int rc = dalloc(&s->G.velG, size_t(s->scap) * 512 * 3);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(s->G.velG, 0, size_t(s->scap) * 512 * 12, s->st));
    s->lcap = s->scap * 512u;
  }
  s->initStep = s->nbOn ? o->nb_init : 0u;
  if (s->nbOn) {
    int rc = ensureNb(s);
    if (rc) return rc;
  }
  uint32_t nc = s->ncTotal >= o->nfv ? s->ncTotal - o->nfv : 0u;
  {  // the per-cell collider lookup only changes when a collider moves (sim.js:191-195)
    bool same = s->haveCols && s->sdfSig == s->sdfSigPrev && !o->no_solid_cache && s->cfPrev.size() == size_t(nc) * 32;
    if (same)
      for (size_t q = 0; q < size_t(nc) * 32; q++)
        if (s->cfPrev[q] != s->cf[q]) { same = false; break; }
    if (!same) {
      s->solidDirty = true;
      s->cfPrev.assign(s->cf.begin(), s->cf.begin() + size_t(nc) * 32);
      s->sdfSigPrev = s->sdfSig;
      s->haveCols = true;
    }
  }
  s->onlyTiles = o->only_tiles != 0;
  if (o->sor_mode == ST_SOR_TUNED) {  // the path st_tune_sor chose (the resident solve until then)
    s->sorMode = s->tuned ? s->tunedMode : 0u;
    s->pipeOverride = s->tuned && s->tunedMode == 1u ? s->tunedPipe : 0u;
  } else {
    s->sorMode = o->sor_mode;
    s->pipeOverride = 0u;
  }
  DParams& P = s->hp;
  std::memset(&P, 0, sizeof P);
  P.n = s->N; P.nx = uint32_t(s->nx); P.ny = uint32_t(s->ny); P.nz = uint32_t(s->nz); P.ncols = nc; P.open = s->open; P.ncell = s->ncell; P.scap = s->scap;
  double omegaDef = std::min(1.9, 2.0 / (1.0 + std::sin(M_PI / std::max(s->nx, std::max(s->ny, s->nz)))));
  P.dt = o->dt; P.h = float(s->hd); P.ih = float(1.0 / s->hd); P.gravity = o->g; P.flip = o->flip; P.rest = s->rest; P.comp = o->comp;
  P.omega = o->omega != 0.0f ? o->omega : float(omegaDef);
  P.x0 = s->x0; P.y0 = s->y0; P.z0 = s->z0; P.vmax = o->vmax != 0.0f ? o->vmax : 14.0f;
  P.margin = float(0.3 * s->dd); P.scale = 65536.0f; P.x1 = s->x1; P.y1 = s->y1; P.z1 = s->z1; P.q0 = float(0.3 * s->hd);
  P.q1 = o->wind[0]; P.q2 = o->wind[1]; P.q3 = o->wind[2]; P.q4 = o->drag; P.q5 = o->foam_tau; P.q6 = o->foam_sens; P.q7 = o->gust;
  P.q8 = float(s->tm); P.q9 = o->no_foam ? 1.0f : 0.0f; P.q10 = o->no_cull ? 1.0f : 0.0f; P.q11 = 1.0f; P.q12 = o->fast_air ? 1.0f : 0.0f;
  if (!s->onlyTiles) s->tm = s->tm + double(o->dt);
  P.sheetMin = o->tear;
  P.band = std::max(2.0f, o->nb_band > 0 ? o->nb_band : 3.0f);
  P.tx = s->tx; P.ty = s->ty; P.tz = s->tz; P.lcap = s->lcap;
  if (!s->onlyTiles) s->seed = s->seed + 1u;
  P.seed = s->seed;
  P.nb = s->nbOn ? 1u : 0u;
  P.nbInit = s->initStep;
  P.nfill = s->initStep ? std::min<uint32_t>(o->nfill, 8u) : 0u;
  P.nfv = s->initStep ? o->nfv : 0u;
  P.ntiles = s->ntiles;
  P.settle = st_smoothstepf(0.1f, 0.5f, P.q8);
  {
    const float sens = std::fmax(P.q6, 0.05f);
    P.impS2 = quietBelow(15.0, P.dt, sens);
    P.turbS2 = quietBelow(double(0.35f), 1.0f, sens);
  }
  {  // whitewater, air field (sim.js:203-206, 221-230)
    const bool on = o->spr_on != 0 && s->sprayCap > 0, haveSp = o->spr_on != 0;
    P.sprRate = on ? o->spr_rate : 0.0f;
    P.sprMin = on ? std::max(0.05f, o->spr_min) : 1.0f;
    P.sprLife = on ? o->spr_life : 1.0f;
    P.mistFrac = on ? o->spr_mist : 0.0f;
    P.sprDrag = on ? o->spr_drag : 1.0f;
    P.mistDrag = on ? o->spr_mdrag : 4.0f;
    P.vs = o->vs != 0.0f ? o->vs : 1.0f;
    P.smax = s->sprayCap;
    // ww = (sp && sp.ww) || {}: without o.spr the defaults hold
    P.wwFoamLife = haveSp && o->ww_foam_life != 0.0f ? o->ww_foam_life : 3.0f;
    P.wwRise = haveSp && o->ww_rise != 0.0f ? o->ww_rise : 0.3f;
    P.wwCrest = haveSp ? o->ww_crest : 1.0f;
    P.wwImpact = haveSp ? o->ww_impact : 0.5f;
    const bool airo = on && o->air_on != 0;
    if (airo && !s->airN) {
      int rc = airAlloc(s);
      if (rc) return rc;
    }
    P.airOn = airo && s->airN ? 1u : 0u;
    P.airCouple = airo ? o->air_couple : 0.0f;
    P.airDecay = airo ? o->air_decay : 0.0f;
    P.airH = s->airN ? float(s->airH) : 1.0f;
    P.aNx = s->airN ? s->aNx : 1u;
    P.aNy = s->airN ? s->aNy : 1u;
    P.aNz = s->airN ? s->aNz : 1u;
    P.airBlur = airo ? o->air_blur : 0.0f;
    P.sprOn = on ? 1u : 0u;
    const bool dr = on && o->drip_on != 0;
    P.ndrip = dr ? s->ndrip : 0u;
    P.dripRate = dr ? o->drip_rate : 0.0f;
    P.dripDry = dr ? o->drip_dry : 1.0f;
    s->airOnStep = P.airOn != 0u;
    s->sprOnStep = on;
    s->dripOnStep = dr && s->ndrip > 0;
  }
  for (int k = 0; k < 16; k++)
    for (int q = 0; q < 4; q++) P.fills[k][q] = o->fills[k * 4 + q];
  {  // ocean (sim.js:207-218): words 57-63 and 66, the waves from word 148
    const bool oc = o->ocean_on != 0;
    P.ocean = oc ? 1u : 0u;
    P.nwave = oc ? o->nwave : 0u;
    P.otime = o->ocean_time;
    P.olevel = oc ? o->ocean_level : 0.0f;
    P.olayer = oc ? o->ocean_layer : 0.0f;
    P.oy0 = s->y0;
    P.odepth = oc ? float(std::max(double(o->ocean_level) - double(s->y0), 1e-3)) : 1.0f;
    P.oease = 1.0f - st_expf(-8.0f * P.dt);
    for (uint32_t k = 0; k < P.nwave; k++)
      for (int q = 0; q < 8; q++) P.waves[k * 2 + q / 4][q % 4] = o->waves[k * 8 + q];
  }
  // opt-in modes (no counterpart in the reference)
  if (o->extrap_layers > ST_EXTRAP_MAX) return fail(ST_ERR_ARG, "st_params.extrap_layers: at most ST_EXTRAP_MAX layers");
  P.rk2 = o->rk2 ? 1u : 0u;
  P.extrap = o->extrap_layers;
  if (o->extrap_layers && !o->only_tiles && (!s->exMask || s->exCells < s->scap * 512u)) {
    HIPCHK(hipStreamSynchronize(s->st));
    dfree(s->exMask);
    s->exCells = 0;
    int rc = dalloc(&s->exMask, size_t(s->scap) * 512u * 2u);
    if (rc) return rc;
    s->exCells = s->scap * 512u;
  }
  s->extrapOnStep = o->only_tiles ? 0u : o->extrap_layers;
  s->iters = std::max(1, int(o->iters));
  {
    // Every halo exchange of a resident solve carries its own tag, taken from
    // this block (so each replay of a captured substep gets new ones): the
    // step reserves 2 iters + 2.  When the counter would wrap, the entries are
    // cleared first so that no old tag can match.
    const uint32_t need = 2u * uint32_t(s->iters) + 2u;
    if (s->sorTag > 0xffffffffu - need) {
      launch_fill_u32(s->st, reinterpret_cast<uint32_t*>(s->G.sorHalo), size_t(s->scap) * 2 * 6 * 32 * 2, 0u);
      s->sorTag = 1;
    }
    P.sorTag = s->sorTag;
    P.iters = uint32_t(s->iters);
    s->sorTag += need;
  }
  Pout = P;
  size_t ncAll = std::min<size_t>(s->ncTotal, ST_MAX_COLS);
  std::memcpy(C.c, s->cf.data(), ncAll * 32 * sizeof(float));
  for (size_t q = 0; q < ncAll; q++) {
    float hq = P.q0 / s->cf[q * 32 + 7];  // q0 / cell size of the collider grid
    C.near[q] = make_float4(-1.0f - hq, hq, 0.0f, 0.0f);
    nearBox(&s->cf[q * 32], -1.0f - hq, hq, C.wlo[q], C.whi[q]);
  }
  return ST_OK;
}

// sets and uploads the step's parameter block and collider records
int beginStep(Sim* s, const st_params* o) {
  // step slot: wait until the ring entry's last copy finished, fill, copy
  uint32_t k = s->ringPos % RING;
  if (s->evUsed[k]) HIPCHK(hipEventSynchronize(s->ev[k]));
  StepSlot& hs = s->hSlots[k];
  int rc = prepStep(s, o, hs.P, hs.C);
  if (rc) return rc;
  s->ringPos++;
  HIPCHK(hipMemcpyAsync(&s->dSlots[k], &hs, sizeof(StepSlot), hipMemcpyHostToDevice, s->st));
  if (!s->evUsed[k]) {
    HIPCHK(hipEventCreateWithFlags(&s->ev[k], hipEventDisableTiming));
    s->evUsed[k] = true;
  }
  HIPCHK(hipEventRecord(s->ev[k], s->st));
  s->cur = &s->dSlots[k];
  s->begun = true;
  return ST_OK;
}

int runStage(Sim* s, int stage) {
  if (!s->begun) return fail(ST_ERR_STATE, "st_stage: no step begun");
  LaunchCtx L = s->ctx();
  hipStream_t st = s->st;
  bool full = !s->onlyTiles;
  switch (stage) {
    case ST_STAGE_SORT: break;
    case ST_STAGE_BLOCKS: launch_blocks(st, L); s->binned = false; break;
    case ST_STAGE_CLEAR: if (full) launch_clear(st, L); break;
    case ST_STAGE_P2G: if (full) launch_p2g(st, L); break;
    case ST_STAGE_NARROWBAND: if (full && s->nbOn) launch_narrowband(st, L); break;
    case ST_STAGE_CLASSIFY:
      if (full) {
        launch_classify(st, L, s->solidDirty);
        s->solidDirty = false;
      }
      break;
    case ST_STAGE_NORMFACES: if (full) launch_norm_faces(st, L); break;
    case ST_STAGE_DIVERGENCE: if (full) launch_divergence(st, L); break;
    case ST_STAGE_PRESSURE:
      if (full) {
        L.sorSeq = ++s->sorSeq;
        if (launch_pressure(st, L, s->iters) != hipSuccess) s->coopGrid = 0;  // per-iteration launches from now on
      }
      break;
    case ST_STAGE_PROJECT: if (full) launch_project(st, L); break;
    case ST_STAGE_G2P:
      if (full) {
        launch_g2p(st, L);
        s->binned = L.binNext && L.n != 0;
        if (L.n) {  // g2p wrote the new state into the other buffer
          std::swap(s->G.pos, s->G.posB);
          std::swap(s->G.vel, s->G.velB);
          std::swap(s->G.pid, s->G.pidB);
        }
      }
      break;
    case ST_STAGE_SPRAY:
      if (full && s->sprayCap) launch_spray(st, L, s->airN, s->airOnStep, s->sprOnStep, s->dripOnStep ? s->ndrip : 0u);
      break;
    case ST_STAGE_RESEED:
      if (full && s->nbOn) {
        launch_reseed(st, L, s->nb);
        s->binned = false;
      }
      break;
    default: return fail(ST_ERR_ARG, "st_stage: unknown stage");
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(ST_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return ST_OK;
}

// Graph replay of a plain substep (st_step_n with use_graph).  The step's
// parameter block and collider records go to one fixed device slot, which the
// captured kernels read; the first substep of a kind is captured (stream
// capture of every stage) and later ones replay it.  A graphed substep solves
// with one launch per SOR iteration (the same bits).  A captured cooperative
// launch does replay, alone with the right bits, but with a second sim
// stepping on its own stream at the same time the replays gave other bits
// than the eager substeps (direct cooperative launches kept them): a graph
// node is not scheduled as a cooperative launch, whose workgroups must all be
// resident at once.
// The host-side effects of the stages (the particle buffers swapped, the bins
// of the next substep made, the collider lookup rebuilt) are applied again on
// each replay.
void dropGraphs(Sim* s) {
  if (s->graphs.empty()) return;
  (void)hipStreamSynchronize(s->st);
  for (StepGraph& g : s->graphs) (void)hipGraphExecDestroy(g.exec);
  s->graphs.clear();
}

int graphStep(Sim* s, const st_params* o) {
  if (o->nb_on || o->ocean_on || o->only_tiles || o->nb_init)
    return fail(ST_ERR_ARG, "st_step_n: graph replay covers plain substeps (no narrow band or ocean, no tile-only step)");
  if (s->batch) return fail(ST_ERR_STATE, "st_step_n: graph replay of a tank in a batch (use st_batch_step_n)");
  if (!s->dGraphSlot) {
    int rc = dalloc(&s->dGraphSlot, 1);
    if (rc) return rc;
  }
  uint32_t k = s->ringPos % RING;
  if (s->evUsed[k]) HIPCHK(hipEventSynchronize(s->ev[k]));
  StepSlot& hs = s->hSlots[k];
  st_params q = *o;
  q.sor_mode = 1;
  int rc = prepStep(s, &q, hs.P, hs.C);
  if (rc) return rc;
  if (o->sor_mode == ST_SOR_TUNED && s->tuned && s->tunedMode == 1u) s->pipeOverride = s->tunedPipe;
  s->ringPos++;
  HIPCHK(hipMemcpyAsync(s->dGraphSlot, &hs, sizeof(StepSlot), hipMemcpyHostToDevice, s->st));
  if (!s->evUsed[k]) {
    HIPCHK(hipEventCreateWithFlags(&s->ev[k], hipEventDisableTiming));
    s->evUsed[k] = true;
  }
  HIPCHK(hipEventRecord(s->ev[k], s->st));
  s->cur = s->dGraphSlot;
  s->begun = true;
  const uint32_t pipe = s->ctx().pipeGrid;
  for (StepGraph& g : s->graphs) {
    if (std::memcmp(&g.G, &s->G, sizeof(DGrid)) != 0 || g.binned != s->binned || g.solid != s->solidDirty ||
        g.iters != s->iters || g.N != s->N || g.scap != s->scap || g.pipe != pipe || g.airOn != s->airOnStep ||
        g.sprOn != s->sprOnStep || g.dripOn != s->dripOnStep || g.ndrip != s->ndrip || g.airN != s->airN ||
        g.sprayCap != s->sprayCap || g.extrap != s->extrapOnStep)
      continue;
    HIPCHK(hipGraphLaunch(g.exec, s->st));
    // what runStage does on the host for these stages
    s->solidDirty = false;
    s->sorSeq++;
    s->binned = true;
    std::swap(s->G.pos, s->G.posB);
    std::swap(s->G.vel, s->G.velB);
    std::swap(s->G.pid, s->G.pidB);
    return ST_OK;
  }
  StepGraph e;
  e.G = s->G;
  e.binned = s->binned;
  e.solid = s->solidDirty;
  e.iters = s->iters;
  e.N = s->N;
  e.scap = s->scap;
  e.pipe = pipe;
  e.airOn = s->airOnStep;
  e.sprOn = s->sprOnStep;
  e.dripOn = s->dripOnStep;
  e.ndrip = s->ndrip;
  e.airN = s->airN;
  e.sprayCap = s->sprayCap;
  e.extrap = s->extrapOnStep;
  e.exec = nullptr;
  HIPCHK(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
  for (int st = 0; st < ST_STAGE_COUNT && !rc; st++) rc = runStage(s, st);
  hipGraph_t g = nullptr;
  hipError_t ce = hipStreamEndCapture(s->st, &g);
  if (rc || ce != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return rc ? rc : fail(ST_ERR_HIP, std::string("st_step_n: stream capture: ") + hipGetErrorString(ce));
  }
  hipError_t ie = hipGraphInstantiate(&e.exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ie != hipSuccess) return fail(ST_ERR_HIP, std::string("st_step_n: hipGraphInstantiate: ") + hipGetErrorString(ie));
  HIPCHK(hipGraphLaunch(e.exec, s->st));
  if (s->graphs.size() >= GRAPH_CACHE) {
    HIPCHK(hipStreamSynchronize(s->st));
    (void)hipGraphExecDestroy(s->graphs.front().exec);
    s->graphs.erase(s->graphs.begin());
  }
  s->graphs.push_back(e);
  return ST_OK;
}

// Batched tanks (no counterpart in the reference, which steps one FlipSim at a
// time): every stage of a substep for all tanks in one launch per kernel over
// the concatenated workgroup ranges of the tanks.  The tanks keep their own
// state; while in the batch they share the batch's stream.
const int BRING = 8;
struct Batch {
  int dev = 0;
  hipStream_t st = nullptr;
  std::vector<Sim*> tanks;
  uint32_t sorResident = 2048;   // SOR workgroups of a batched iteration, over all tanks
  uint32_t coopGrid = 0;         // resident workgroups of the one-launch batched SOR (0: launches per iteration)
  uint32_t* dPlan = nullptr;     // its per-tank workgroup plan (3 nt + 1 words)
  uint32_t lastResident = 0;     // the last substep's SOR ran in the one launch
  // ring of step slots: tank records, parameter blocks, collider records, block
  // tables (host pinned and device), each reusable once its upload is done
  size_t slotBytes = 0, offP = 0, offC = 0, offTab = 0;
  uint32_t itersCap = 0;
  uint8_t* hRing = nullptr;
  uint8_t* dRing = nullptr;
  hipEvent_t ev[BRING];
  bool evUsed[BRING];
  uint32_t ringPos = 0;
  std::vector<BTankPlan> plan;
  // graph replay (st_batch_step_n with use_graph): a fixed device slot the
  // captured kernels read, and the captured substeps by their block tables
  uint8_t* dGraph = nullptr;
  struct Captured {
    std::vector<uint32_t> tab;
    hipGraphExec_t exec;
  };
  std::vector<Captured> graphs;
};

size_t align256(size_t x) { return (x + 255u) & ~size_t(255u); }

void freeRing(Batch* b) {
  for (int k = 0; k < BRING; k++)
    if (b->evUsed[k]) {
      (void)hipEventSynchronize(b->ev[k]);
      (void)hipEventDestroy(b->ev[k]);
      b->evUsed[k] = false;
    }
  if (b->hRing) (void)hipHostFree(b->hRing);
  dfree(b->dRing);
  if (!b->graphs.empty() || b->dGraph) (void)hipStreamSynchronize(b->st);
  for (auto& g : b->graphs) (void)hipGraphExecDestroy(g.exec);
  b->graphs.clear();
  dfree(b->dGraph);
  b->hRing = nullptr;
  b->itersCap = 0;
}

// ring slots sized for nt tanks and up to `iters` SOR iterations
int ensureRing(Batch* b, uint32_t iters) {
  if (b->hRing && iters <= b->itersCap) return ST_OK;
  HIPCHK(hipStreamSynchronize(b->st));
  freeRing(b);
  const size_t nt = b->tanks.size();
  b->offP = align256(nt * sizeof(BTank));
  b->offC = b->offP + align256(nt * sizeof(DParams));
  b->offTab = b->offC + align256(nt * sizeof(DCols));
  b->slotBytes = align256(b->offTab + batch_table_words(uint32_t(nt), iters) * 4u);
  hipError_t e = hipHostMalloc((void**)&b->hRing, b->slotBytes * BRING, hipHostMallocDefault);
  if (e != hipSuccess) {
    b->hRing = nullptr;
    return fail(ST_ERR_NOMEM, "hipHostMalloc failed");
  }
  std::memset(b->hRing, 0, b->slotBytes * BRING);
  int rc = dalloc(&b->dRing, b->slotBytes * BRING);
  if (rc) return rc;
  b->itersCap = iters;
  return ST_OK;
}

// A copy of the state a substep starts from (particles, tile tables, pool
// pressure and cell types, host scalars), to re-run substeps from it.
struct Snap {
  float4* pos = nullptr;
  float4* vel = nullptr;
  uint32_t* pid = nullptr;
  uint32_t *tmap = nullptr, *slotTile = nullptr, *fstack = nullptr, *slotList = nullptr, *ctl = nullptr, *ctype = nullptr;
  float* pres = nullptr;
  uint32_t N = 0, scap = 0, seed = 0;
  double tm = 0;
};
void freeSnap(Snap* q) {
  dfree(q->pos); dfree(q->vel); dfree(q->pid); dfree(q->tmap); dfree(q->slotTile); dfree(q->fstack); dfree(q->slotList);
  dfree(q->ctl); dfree(q->ctype); dfree(q->pres);
  delete q;
}

// Adjoint context: the tapes of recorded substeps (a stack) and the gradients.
struct Adj {
  Sim* s = nullptr;
  uint32_t maxSteps = 0, depth = 0, scap = 0;
  std::vector<AdjTape> tapes;
  std::vector<uint64_t> sdfSig;   // per tape: the collider grids it was recorded with
  std::vector<const float*> sdfAt;
  std::vector<uint32_t> ncols;    // per tape: colliders of its substep
  uint32_t lastCols = 0;          // colliders of the substep reversed last
  uint32_t coopGrid = 0;          // workgroups of the one-launch transposed solve (0: one launch per sweep)
  bool oneLaunch = false;         // st_adj_set_sor_path: use the one launch when coopGrid > 0
  bool lastOne = false;           // the last reversed substep's solve ran as one launch
  AdjGrad G{};
};
void freeTape(AdjTape& T) {
  dfree(T.pos); dfree(T.vel); dfree(T.pid); dfree(T.binStart); dfree(T.binIdx); dfree(T.tmap0); dfree(T.tmap); dfree(T.slotTile); dfree(T.slotList); dfree(T.ctl);
  dfree(T.ctype); dfree(T.sten); dfree(T.accWraw); dfree(T.accVfin); dfree(T.accWfin); dfree(T.P); dfree(T.C);
}
void freeAdj(Adj* a) {
  for (AdjTape& T : a->tapes) freeTape(T);
  AdjGrad& G = a->G;
  dfree(G.gPos); dfree(G.gVel); dfree(G.gPos2); dfree(G.gVel2); dfree(G.gFN); dfree(G.gFO); dfree(G.gPres); dfree(G.gPres2);
  dfree(G.gRhs); dfree(G.carry); dfree(G.sorNb); dfree(G.sorCoef); dfree(G.gCol); dfree(G.bar);
  delete a;
}
int allocTape(Sim* s, AdjTape& T) {
  size_t M = s->maxn, C = size_t(s->scap) * 512u;
  int rc;
  if ((rc = dalloc(&T.pos, M)) || (rc = dalloc(&T.vel, M)) || (rc = dalloc(&T.pid, M)) || (rc = dalloc(&T.tmap0, s->ntiles)) ||
      (rc = dalloc(&T.binStart, size_t(s->ntiles) + 2)) || (rc = dalloc(&T.binIdx, M)) ||
      (rc = dalloc(&T.tmap, s->ntiles)) || (rc = dalloc(&T.slotTile, s->scap)) || (rc = dalloc(&T.slotList, s->scap)) ||
      (rc = dalloc(&T.ctl, 16)) || (rc = dalloc(&T.ctype, C)) || (rc = dalloc(&T.sten, C)) || (rc = dalloc(&T.accWraw, C * 3)) ||
      (rc = dalloc(&T.accVfin, C * 3)) || (rc = dalloc(&T.accWfin, C * 3)) || (rc = dalloc(&T.P, 1)) || (rc = dalloc(&T.C, 1)))
    return rc;
  return ST_OK;
}

#define D2D(dst, src, n) HIPCHK(hipMemcpyAsync((dst), (src), (n) * sizeof(*(src)), hipMemcpyDeviceToDevice, s->st))

}  // namespace

extern "C" {

const char* st_last_error(void) { return g_err.c_str(); }
const char* st_version(void) { return "splashtank-mi355x 0.1 (gfx950)"; }

void st_params_default(st_params* p) {
  std::memset(p, 0, sizeof *p);
  p->dt = 1.0f / 60.0f;
  p->g = 9.81f;
  p->flip = 0.95f;
  p->comp = 6.0f;
  p->omega = 0.0f;
  p->vmax = 14.0f;
  p->tear = 0.0f;
  p->iters = 30;
  p->foam_tau = 1.2f;
  p->foam_sens = 1.0f;
  p->nb_band = 3.0f;
  p->ww_crest = 1.0f;
  p->ww_impact = 0.5f;
  p->air_blur = 0.35f;
}

int st_create(const st_opts* o, void** out) {
  if (!o || !out) return fail(ST_ERR_ARG, "st_create: null argument");
  *out = nullptr;
  if (!(o->d > 0) || o->max_n == 0) return fail(ST_ERR_ARG, "st_create: d must be > 0 and max_n > 0");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (o->device < 0 || o->device >= ndev) return fail(ST_ERR_ARG, "st_create: no such HIP device");
  HIPCHK(hipSetDevice(o->device));
  Sim* s = new Sim();
  for (int k = 0; k < RING; k++) s->evUsed[k] = false;
  s->dev = o->device;
  {
    int cus = 0;
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, o->device));
    s->sorResident = sor_resident_grid(uint32_t(cus));
    s->coopGrid = sor_coop_grid(uint32_t(cus));
  }
  double rt = o->ratio > 0 ? o->ratio : 2.0;
  rt = std::max(1.4, std::min(4.0, rt));
  s->ratio = float(rt);
  s->dd = o->d;
  s->hd = rt * o->d;
  s->d = float(o->d);
  s->h = float(s->hd);
  s->rest = float(rt * rt * rt);
  double sz[3] = {o->size[0] > 0 ? o->size[0] : 3.0, o->size[1] > 0 ? o->size[1] : 2.6, o->size[2] > 0 ? o->size[2] : 3.0};
  s->nx = std::max(4, (int)std::floor(sz[0] / s->hd + 0.5));
  s->ny = std::max(4, (int)std::floor(sz[1] / s->hd + 0.5));
  s->nz = std::max(4, (int)std::floor(sz[2] / s->hd + 0.5));
  if (double(s->nx) * s->ny * s->nz >= 4294967296.0 || uint64_t((s->nx + 8) / 8) * uint64_t((s->ny + 8) / 8) * uint64_t((s->nz + 8) / 8) >= (1ull << 24)) {
    delete s;
    return fail(ST_ERR_ARG, "st_create: grid too large (cells below 2^32 and 8^3 tiles below 2^24)");
  }
  double W = s->nx * s->hd, H = s->ny * s->hd, D = s->nz * s->hd;
  s->x0 = float(-W / 2); s->x1 = float(W / 2); s->y0 = 0.0f; s->y1 = float(H); s->z0 = float(-D / 2); s->z1 = float(D / 2);
  s->open = o->open;
  s->maxn = o->max_n;
  s->ncell = uint32_t(s->nx) * s->ny * s->nz;
  s->tx = (s->nx + 1 + 7) / 8; s->ty = (s->ny + 1 + 7) / 8; s->tz = (s->nz + 1 + 7) / 8;
  s->ntiles = s->tx * s->ty * s->tz;
  s->scapMax = s->ntiles + 1;
  s->scap = o->slots ? std::max<uint32_t>(2, std::min(s->scapMax, o->slots)) : s->scapMax;
  s->lcap = std::min<uint32_t>(s->scap * 512u, s->maxn);
  int rc = ST_OK;
  hipError_t e = hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking);
  if (e != hipSuccess) { destroySim(s); return fail(ST_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)); }
  size_t M = s->maxn;
  uint32_t nbin = s->ntiles + 1;
  if ((rc = dalloc(&s->G.pos, M)) || (rc = dalloc(&s->G.vel, M)) || (rc = dalloc(&s->G.pid, M)) || (rc = dalloc(&s->G.scratch4, M)) ||
      (rc = dalloc(&s->G.posB, M)) || (rc = dalloc(&s->G.velB, M)) || (rc = dalloc(&s->G.pidB, M)) || (rc = dalloc(&s->G.binKey, M)) || (rc = dalloc(&s->G.binIdx, M)) ||
      (rc = dalloc(&s->G.binCount, nbin)) || (rc = dalloc(&s->G.binStart, size_t(nbin) + 1)) || (rc = dalloc(&s->G.binRank, M)) ||
      (rc = dalloc(&s->G.scanTmp, size_t(nbin + 1023) / 1024 + 1)) ||
      (rc = dalloc(&s->G.tmap, s->ntiles)) || (rc = dalloc(&s->G.need, s->ntiles)) || (rc = dalloc(&s->G.newList, size_t(s->ntiles) * 2)) ||
      (rc = dalloc(&s->G.wg, size_t((s->ntiles + 255) / 256) * 3 + 3)) || (rc = dalloc(&s->G.ctl, 16)) || (rc = allocSlots(s, s->scap)) ||
      (rc = dalloc(&s->G.part, 2 * 8192 + 64)) || (rc = dalloc(&s->G.foamTab, FOAM_TAB)) ||
      (rc = dalloc(&s->G.mortonTiles, s->ntiles)) || (rc = dalloc(&s->G.coopFlags, s->ntiles)) ||
      (rc = dalloc(&s->G.coopOffs, size_t(s->ntiles) + 1)) || (rc = dalloc(&s->G.sorBar, COOP_MAX_WG + 4)) ||
      (rc = dalloc(&s->dSlots, RING))) {
    destroySim(s);
    return rc;
  }
  e = hipHostMalloc((void**)&s->hSlots, sizeof(StepSlot) * RING, hipHostMallocDefault);
  if (e != hipSuccess) { destroySim(s); return fail(ST_ERR_NOMEM, "hipHostMalloc failed"); }
  std::memset(s->hSlots, 0, sizeof(StepSlot) * RING);
  // empty tile map, every slot but the spare slot 0 on the free stack (sim.js:94-98)
  std::vector<uint32_t> host(std::max<size_t>(M, std::max<size_t>(s->ntiles, s->scap)));
  for (size_t k = 0; k < M; k++) host[k] = uint32_t(k);
  // ids in both particle buffers: slots past the live count keep their own id,
  // whichever buffer is current, so appended particles get id = index
  if ((e = hipMemcpy(s->G.pid, host.data(), M * 4, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(s->G.pidB, host.data(), M * 4, hipMemcpyHostToDevice)) != hipSuccess) {
    destroySim(s);
    return fail(ST_ERR_HIP, "pid upload failed");
  }
  std::vector<uint32_t> st(s->scap, 0u);
  for (uint32_t k = 0; k + 1 < s->scap; k++) st[k] = s->scap - 1 - k;
  std::vector<uint32_t> tm(s->ntiles, ST_NONE);
  std::vector<uint32_t> mort(s->ntiles);
  morton_tiles(s->tx, s->ty, s->tz, mort.data());
  uint32_t ctl[16] = {0, 0, s->scap - 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (hipMemcpy(s->G.fstack, st.data(), size_t(s->scap) * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(s->G.tmap, tm.data(), size_t(s->ntiles) * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(s->G.mortonTiles, mort.data(), size_t(s->ntiles) * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(s->G.sorBar, 0, (COOP_MAX_WG + 4) * 4) != hipSuccess ||
      hipMemcpy(s->G.ctl, ctl, 64, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(s->G.pos, 0, M * 16) != hipSuccess || hipMemset(s->G.vel, 0, M * 16) != hipSuccess ||
      hipMemset(s->G.need, 0, size_t(s->ntiles) * 4) != hipSuccess) {
    destroySim(s);
    return fail(ST_ERR_HIP, "st_create: initial upload failed");
  }
  if ((rc = zeroSlots(s, 0, s->scap))) { destroySim(s); return rc; }
  e = hipStreamSynchronize(s->st);
  if (e != hipSuccess) { destroySim(s); return fail(ST_ERR_HIP, "st_create: sync failed"); }
  *out = s;
  return ST_OK;
}

int st_destroy(void* p) {
  if (!p) return fail(ST_ERR_ARG, "st_destroy: null sim");
  if (getSim(p)->batch) return fail(ST_ERR_STATE, "st_destroy: the tank is in a batch (destroy the batch first)");
  destroySim(getSim(p));
  return ST_OK;
}

int st_get_dims(void* p, st_dims* o) {
  Sim* s = getSim(p);
  if (!s || !o) return fail(ST_ERR_ARG, "st_get_dims: null argument");
  o->d = s->d; o->h = s->h; o->ratio = s->ratio; o->rest = s->rest;
  o->x0 = s->x0; o->y0 = s->y0; o->z0 = s->z0; o->x1 = s->x1; o->y1 = s->y1; o->z1 = s->z1;
  o->nx = s->nx; o->ny = s->ny; o->nz = s->nz; o->tx = int32_t(s->tx); o->ty = int32_t(s->ty); o->tz = int32_t(s->tz);
  o->ntiles = s->ntiles; o->scap = s->scap; o->scap_max = s->scapMax; o->max_n = s->maxn; o->ncell = s->ncell;
  return ST_OK;
}

int st_upload_particles(void* p, uint32_t first, uint32_t count, const float* pos4, const float* vel4) {
  Sim* s = getSim(p);
  if (!s || (count && (!pos4 || !vel4))) return fail(ST_ERR_ARG, "st_upload_particles: null argument");
  if (size_t(first) + count > s->maxn) return fail(ST_ERR_ARG, "st_upload_particles: beyond max_n");
  if (!count) return ST_OK;
  HIPCHK(hipSetDevice(s->dev));
  HIPCHK(hipMemcpyAsync(s->G.pos + first, pos4, size_t(count) * 16, hipMemcpyHostToDevice, s->st));
  HIPCHK(hipMemcpyAsync(s->G.vel + first, vel4, size_t(count) * 16, hipMemcpyHostToDevice, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  s->N = std::max(s->N, first + count);
  s->binned = false;
  return ST_OK;
}

int st_set_count(void* p, uint32_t n) {
  Sim* s = getSim(p);
  if (!s) return fail(ST_ERR_ARG, "st_set_count: null sim");
  if (n > s->maxn) return fail(ST_ERR_ARG, "st_set_count: beyond max_n");
  s->N = n;
  s->binned = false;
  return ST_OK;
}

int st_get_count(void* p, uint32_t* n) {
  Sim* s = getSim(p);
  if (!s || !n) return fail(ST_ERR_ARG, "st_get_count: null argument");
  *n = s->N;
  return ST_OK;
}

int st_clear(void* p) {
  Sim* s = getSim(p);
  if (!s) return fail(ST_ERR_ARG, "st_clear: null sim");
  HIPCHK(hipSetDevice(s->dev));
  HIPCHK(hipStreamSynchronize(s->st));
  s->N = 0;
  s->tm = 0;
  s->binned = false;
  std::vector<uint32_t> ids(s->maxn);
  for (size_t k = 0; k < ids.size(); k++) ids[k] = uint32_t(k);
  HIPCHK(hipMemcpy(s->G.pid, ids.data(), ids.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(s->G.pidB, ids.data(), ids.size() * 4, hipMemcpyHostToDevice));
  if (s->sprayCap) {  // clear() (sim.js:543-544): no whitewater, the ring starts again, calm air
    HIPCHK(hipMemset(s->G.sPos, 0, size_t(s->sprayCap) * 16));
    HIPCHK(hipMemset(s->G.sctl, 0, 16));
  }
  if (s->airN) HIPCHK(hipMemset(s->G.air, 0, size_t(s->airN) * 16));
  return ST_OK;
}

int st_set_colliders(void* p, const float* cf32, uint32_t nc, const float* sdf, size_t nsdf, uint64_t sig) {
  Sim* s = getSim(p);
  if (!s || (nc && !cf32) || (nsdf && !sdf)) return fail(ST_ERR_ARG, "st_set_colliders: null argument");
  if (nc > ST_MAX_COLS) return fail(ST_ERR_ARG, "st_set_colliders: at most 32 colliders");
  HIPCHK(hipSetDevice(s->dev));
  // every grid index a collider record can reach must lie inside the uploaded grids
  for (uint32_t k = 0; k < nc; k++) {
    const float* c = cf32 + size_t(k) * 32;
    double off = c[0], gx = c[1], gy = c[2], gz = c[3];
    if (!(off >= 0) || !(gx >= 2) || !(gy >= 2) || !(gz >= 2) || !(c[7] > 0) || off + gx * gy * gz > double(nsdf))
      return fail(ST_ERR_ARG, "st_set_colliders: collider " + std::to_string(k) + " grid outside the sdf data");
  }
  s->cf.assign(cf32, cf32 + size_t(nc) * 32);
  s->ncTotal = nc;
  if (sig != s->sdfSig || nsdf != s->sdfN) {
    HIPCHK(hipStreamSynchronize(s->st));
    if (nsdf > s->sdfCap) {
      dfree(s->dSdf);
      size_t cap = std::max<size_t>(nsdf + nsdf / 3, 16);
      int rc = dalloc(&s->dSdf, cap);
      if (rc) return rc;
      s->sdfCap = cap;
    }
    if (nsdf) HIPCHK(hipMemcpy(s->dSdf, sdf, nsdf * 4, hipMemcpyHostToDevice));
    s->sdfN = nsdf;
    s->sdfSig = sig;
  }
  if (!s->dSdf) {
    int rc = dalloc(&s->dSdf, 16);
    if (rc) return rc;
    s->sdfCap = 16;
  }
  s->G.sdf = s->dSdf;
  return ST_OK;
}

int st_begin_step(void* p, const st_params* prm) {
  Sim* s = getSim(p);
  if (!s || !prm) return fail(ST_ERR_ARG, "st_begin_step: null argument");
  HIPCHK(hipSetDevice(s->dev));
  if (!s->dSdf) {
    int rc = st_set_colliders(p, nullptr, 0, nullptr, 0, 0);
    if (rc) return rc;
  }
  return beginStep(s, prm);
}

int st_stage(void* p, int32_t stage) {
  Sim* s = getSim(p);
  if (!s) return fail(ST_ERR_ARG, "st_stage: null sim");
  return runStage(s, stage);
}

int st_step(void* p, const st_params* prm) {
  int rc = st_begin_step(p, prm);
  if (rc) return rc;
  Sim* s = getSim(p);
  for (int k = 0; k < ST_STAGE_COUNT; k++)
    if ((rc = runStage(s, k))) return rc;
  return ST_OK;
}

int st_step_n(void* p, const st_params* prm, int32_t nsteps, int32_t use_graph) {
  Sim* s = getSim(p);
  if (!s || !prm) return fail(ST_ERR_ARG, "st_step_n: null argument");
  if (use_graph) HIPCHK(hipSetDevice(s->dev));
  for (int32_t k = 0; k < nsteps; k++) {
    int rc = use_graph ? graphStep(s, prm) : st_step(p, prm);
    if (rc) return rc;
  }
  return ST_OK;
}

int st_time_stages(void* p, const st_params* prm, float ms[ST_STAGE_COUNT]) {
  Sim* s = getSim(p);
  if (!s || !prm || !ms) return fail(ST_ERR_ARG, "st_time_stages: null argument");
  int rc = st_begin_step(p, prm);
  if (rc) return rc;
  hipEvent_t e[ST_STAGE_COUNT + 1];
  for (int k = 0; k <= ST_STAGE_COUNT; k++) HIPCHK(hipEventCreate(&e[k]));
  HIPCHK(hipEventRecord(e[0], s->st));
  for (int k = 0; k < ST_STAGE_COUNT; k++) {
    if ((rc = runStage(s, k))) break;
    HIPCHK(hipEventRecord(e[k + 1], s->st));
  }
  if (!rc) {
    HIPCHK(hipEventSynchronize(e[ST_STAGE_COUNT]));
    for (int k = 0; k < ST_STAGE_COUNT; k++) {
      float t = 0.0f;
      HIPCHK(hipEventElapsedTime(&t, e[k], e[k + 1]));
      ms[k] = t;
    }
  }
  for (int k = 0; k <= ST_STAGE_COUNT; k++) (void)hipEventDestroy(e[k]);
  return rc;
}

int st_read_particles(void* p, float* dst4, uint32_t n) {
  Sim* s = getSim(p);
  if (!s || (n && !dst4)) return fail(ST_ERR_ARG, "st_read_particles: null argument");
  n = std::min(n, s->N);
  if (!n) return ST_OK;
  HIPCHK(hipSetDevice(s->dev));
  launch_unsort(s->st, s->G.pos, s->G.pid, s->G.scratch4, n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(dst4, s->G.scratch4, size_t(n) * 16, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return ST_OK;
}

int st_read_ctl(void* p, uint32_t ctl[16]) {
  Sim* s = getSim(p);
  if (!s || !ctl) return fail(ST_ERR_ARG, "st_read_ctl: null argument");
  HIPCHK(hipSetDevice(s->dev));
  HIPCHK(hipMemcpyAsync(ctl, s->G.ctl, 64, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return ST_OK;
}

int st_grow(void* p, uint32_t n) {
  Sim* s = getSim(p);
  if (!s) return fail(ST_ERR_ARG, "st_grow: null sim");
  n = std::min(s->scapMax, std::max(n, s->scap + 1));
  if (n <= s->scap) return ST_OK;
  HIPCHK(hipSetDevice(s->dev));
  HIPCHK(hipStreamSynchronize(s->st));
  dropGraphs(s);
  uint32_t old = s->scap;
  DGrid keep = s->G;
  s->G.cnt = nullptr; s->G.ctype = nullptr; s->G.pres = nullptr; s->G.pres2 = nullptr; s->G.rhs = nullptr; s->G.sten = nullptr;
  s->G.accV = nullptr; s->G.accW = nullptr; s->G.p2gHalo = nullptr; s->G.velG = nullptr; s->G.slotTile = nullptr; s->G.fstack = nullptr;
  s->G.slotList = nullptr; s->G.nbr = nullptr; s->G.nbrL = nullptr; s->G.coopList = nullptr; s->G.sorHalo = nullptr;
  int rc = allocSlots(s, n);
  if (rc) { freeSlots(s); s->G = keep; return rc; }
  // the wg scratch and grid sizes follow the pool
  size_t C = size_t(old) * 512;
  HIPCHK(hipMemcpy(s->G.cnt, keep.cnt, C * 4, hipMemcpyDeviceToDevice));
  HIPCHK(hipMemcpy(s->G.ctype, keep.ctype, C * 4, hipMemcpyDeviceToDevice));
  HIPCHK(hipMemcpy(s->G.pres, keep.pres, C * 4, hipMemcpyDeviceToDevice));
  HIPCHK(hipMemcpy(s->G.pres2, keep.pres2, C * 4, hipMemcpyDeviceToDevice));
  HIPCHK(hipMemcpy(s->G.rhs, keep.rhs, C * 4, hipMemcpyDeviceToDevice));
  HIPCHK(hipMemcpy(s->G.sten, keep.sten, C, hipMemcpyDeviceToDevice));
  HIPCHK(hipMemcpy(s->G.accV, keep.accV, C * 12, hipMemcpyDeviceToDevice));
  HIPCHK(hipMemcpy(s->G.accW, keep.accW, C * 12, hipMemcpyDeviceToDevice));
  if (keep.velG) HIPCHK(hipMemcpy(s->G.velG, keep.velG, C * 12, hipMemcpyDeviceToDevice));
  HIPCHK(hipMemcpy(s->G.slotTile, keep.slotTile, size_t(old) * 4, hipMemcpyDeviceToDevice));
  HIPCHK(hipMemcpy(s->G.fstack, keep.fstack, size_t(old) * 4, hipMemcpyDeviceToDevice));
  HIPCHK(hipMemcpy(s->G.slotList, keep.slotList, size_t(old) * 4, hipMemcpyDeviceToDevice));
  rc = zeroSlots(s, old, n);
  if (rc) return rc;
  // growFree (wgsl-blocks.js:57-60): the new slots go on the free stack
  uint32_t ctl[16];
  HIPCHK(hipMemcpy(ctl, s->G.ctl, 64, hipMemcpyDeviceToHost));
  std::vector<uint32_t> add(n - old);
  for (uint32_t t = 0; t < n - old; t++) add[t] = old + t;
  HIPCHK(hipMemcpy(s->G.fstack + ctl[2], add.data(), add.size() * 4, hipMemcpyHostToDevice));
  ctl[2] += n - old;
  HIPCHK(hipMemcpy(s->G.ctl, ctl, 64, hipMemcpyHostToDevice));
  freeSlotArrays(keep);
  s->scap = n;
  if (!s->velGOn) s->lcap = std::min<uint32_t>(n * 512u, s->maxn);
  else s->lcap = n * 512u;
  return ST_OK;
}

int st_dump_size(void* p, int32_t what, size_t* bytes) {
  Sim* s = getSim(p);
  if (!s || !bytes) return fail(ST_ERR_ARG, "st_dump_size: null argument");
  HIPCHK(hipSetDevice(s->dev));
  HIPCHK(hipStreamSynchronize(s->st));
  std::vector<uint32_t> tm(s->ntiles);
  HIPCHK(hipMemcpy(tm.data(), s->G.tmap, tm.size() * 4, hipMemcpyDeviceToHost));
  size_t nt = 0;
  for (uint32_t v : tm) nt += v != ST_NONE;
  size_t C = 512;
  switch (what) {
    case ST_DUMP_TILES: *bytes = nt * 4; break;
    case ST_DUMP_CNT: case ST_DUMP_CTYPE: case ST_DUMP_PRES: case ST_DUMP_RHS: *bytes = nt * C * 4; break;
    case ST_DUMP_ACCV: case ST_DUMP_ACCW: case ST_DUMP_FACEV: case ST_DUMP_FACEO: case ST_DUMP_VELG: *bytes = nt * C * 12; break;
    case ST_DUMP_POS: case ST_DUMP_VEL: *bytes = size_t(s->N) * 16; break;
    case ST_DUMP_TMAP: *bytes = size_t(s->ntiles) * 4; break;
    default: return fail(ST_ERR_ARG, "st_dump_size: unknown field");
  }
  return ST_OK;
}

int st_dump(void* p, int32_t what, void* dst, size_t bytes) {
  Sim* s = getSim(p);
  if (!s || !dst) return fail(ST_ERR_ARG, "st_dump: null argument");
  size_t need = 0;
  int rc = st_dump_size(p, what, &need);
  if (rc) return rc;
  if (bytes < need) return fail(ST_ERR_ARG, "st_dump: buffer too small");
  std::vector<uint32_t> tm(s->ntiles);
  HIPCHK(hipMemcpy(tm.data(), s->G.tmap, tm.size() * 4, hipMemcpyDeviceToHost));
  std::vector<uint32_t> ts;
  for (uint32_t t = 0; t < s->ntiles; t++)
    if (tm[t] != ST_NONE) ts.push_back(t);
  char* o = (char*)dst;
  auto perCell = [&](const void* src, size_t per) -> int {
    if (!src) {
      std::memset(o, 0, ts.size() * 512 * per * 4);
      return ST_OK;
    }
    for (size_t q = 0; q < ts.size(); q++) {
      size_t sl = tm[ts[q]];
      hipError_t e = hipMemcpy(o + q * 512 * per * 4, (const char*)src + sl * 512 * per * 4, 512 * per * 4, hipMemcpyDeviceToHost);
      if (e != hipSuccess) return fail(ST_ERR_HIP, "st_dump copy failed");
    }
    return ST_OK;
  };
  switch (what) {
    case ST_DUMP_TILES: std::memcpy(o, ts.data(), ts.size() * 4); return ST_OK;
    case ST_DUMP_CNT: return perCell(s->G.cnt, 1);
    case ST_DUMP_CTYPE: return perCell(s->G.ctype, 1);
    case ST_DUMP_PRES: return perCell(s->G.pres, 1);
    case ST_DUMP_RHS: return perCell(s->G.rhs, 1);
    case ST_DUMP_ACCV: case ST_DUMP_FACEV: return perCell(s->G.accV, 3);
    case ST_DUMP_ACCW: case ST_DUMP_FACEO: return perCell(s->G.accW, 3);
    case ST_DUMP_VELG: return perCell(s->G.velG, 3);
    case ST_DUMP_TMAP: std::memcpy(o, tm.data(), tm.size() * 4); return ST_OK;
    case ST_DUMP_POS:
    case ST_DUMP_VEL: {
      if (!s->N) return ST_OK;
      // id order: out[pid[i]] = field[i]
      std::vector<float> f(size_t(s->N) * 4);
      std::vector<uint32_t> id(s->N);
      HIPCHK(hipMemcpy(f.data(), what == ST_DUMP_POS ? (const void*)s->G.pos : (const void*)s->G.vel, f.size() * 4, hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(id.data(), s->G.pid, id.size() * 4, hipMemcpyDeviceToHost));
      float* out = (float*)o;
      for (size_t i = 0; i < s->N; i++) {
        if (id[i] >= s->N) return fail(ST_ERR_STATE, "st_dump: particle id out of range");
        std::memcpy(out + size_t(id[i]) * 4, f.data() + i * 4, 16);
      }
      return ST_OK;
    }
  }
  return fail(ST_ERR_ARG, "st_dump: unknown field");
}

int st_sync(void* p) {
  Sim* s = getSim(p);
  if (!s) return fail(ST_ERR_ARG, "st_sync: null sim");
  HIPCHK(hipSetDevice(s->dev));
  HIPCHK(hipStreamSynchronize(s->st));
  uint32_t timeouts = 0;
  HIPCHK(hipMemcpy(&timeouts, s->G.sorBar + COOP_MAX_WG + 1, 4, hipMemcpyDeviceToHost));
  if (timeouts) {
    // reported once: the counter starts again from zero, and later solves take
    // the per-iteration launches, which have no waits between workgroups
    HIPCHK(hipMemset(s->G.sorBar + COOP_MAX_WG + 1, 0, 4));
    s->coopGrid = 0;
    return fail(ST_ERR_HIP, "resident SOR: a wait between workgroups timed out (pressure of that solve is not valid); "
                            "later solves use the per-iteration path");
  }
  return ST_OK;
}

int st_set_spray_cap(void* p, uint32_t n) {
  Sim* s = getSim(p);
  if (!s) return fail(ST_ERR_ARG, "st_set_spray_cap: null sim");
  if (n > (1u << 28)) return fail(ST_ERR_ARG, "st_set_spray_cap: at most 2^28 whitewater points");
  if (n == s->sprayCap) return ST_OK;
  HIPCHK(hipSetDevice(s->dev));
  HIPCHK(hipStreamSynchronize(s->st));
  dropGraphs(s);
  dfree(s->G.sPos); dfree(s->G.sVel); dfree(s->G.sAux);
  s->sprayCap = 0;
  if (n == 0) return ST_OK;
  int rc;
  if ((rc = dalloc(&s->G.sPos, n)) || (rc = dalloc(&s->G.sVel, n)) || (rc = dalloc(&s->G.sAux, n))) return rc;
  HIPCHK(hipMemset(s->G.sPos, 0, size_t(n) * 16));
  HIPCHK(hipMemset(s->G.sVel, 0, size_t(n) * 16));
  HIPCHK(hipMemset(s->G.sAux, 0, size_t(n) * 16));
  if (!s->G.sctl) {  // the ring counter outlives the pool, as the reference's bSctl does
    if ((rc = dalloc(&s->G.sctl, 4))) return rc;
    HIPCHK(hipMemset(s->G.sctl, 0, 16));
    const size_t M = size_t(s->maxn) + 1;
    if ((rc = dalloc(&s->G.sprCnt, M)) || (rc = dalloc(&s->G.sprOff, M + 1)) || (rc = dalloc(&s->G.sprSums, M / 1024 + 2))) return rc;
  }
  s->sprayCap = n;
  return ST_OK;
}

int st_set_drips(void* p, const float* samples, uint32_t n) {
  Sim* s = getSim(p);
  if (!s || (n && !samples)) return fail(ST_ERR_ARG, "st_set_drips: null argument");
  HIPCHK(hipSetDevice(s->dev));
  if (n > s->dripCap) {
    HIPCHK(hipStreamSynchronize(s->st));
    dropGraphs(s);
    dfree(s->G.drip); dfree(s->G.dripFlag); dfree(s->G.dripOff); dfree(s->G.dripSums);
    s->dripCap = 0;
    int rc;
    if ((rc = dalloc(&s->G.drip, size_t(n) * 2)) || (rc = dalloc(&s->G.dripFlag, n)) || (rc = dalloc(&s->G.dripOff, size_t(n) + 1)) ||
        (rc = dalloc(&s->G.dripSums, size_t(n) / 1024 + 2)))
      return rc;
    s->dripCap = n;
  }
  if (n) HIPCHK(hipMemcpyAsync(s->G.drip, samples, size_t(n) * 32, hipMemcpyHostToDevice, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  s->ndrip = n;
  return ST_OK;
}

int st_read_spray(void* p, float* pos4, float* vel4, uint32_t room, uint32_t* n) {
  Sim* s = getSim(p);
  if (!s || !n) return fail(ST_ERR_ARG, "st_read_spray: null argument");
  *n = 0;
  if (!s->sprayCap) return ST_OK;
  HIPCHK(hipSetDevice(s->dev));
  uint32_t used = 0;
  HIPCHK(hipMemcpyAsync(&used, s->G.sctl, 4, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  uint32_t k = std::max(1u, std::min(s->sprayCap, used));
  if (k > room || !pos4 || !vel4) return fail(ST_ERR_ARG, "st_read_spray: output arrays too small");
  HIPCHK(hipMemcpyAsync(pos4, s->G.sPos, size_t(k) * 16, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipMemcpyAsync(vel4, s->G.sVel, size_t(k) * 16, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  *n = k;
  return ST_OK;
}

static_assert(DEEP_MAX == ST_DEEP_MAX, "one deep point limit for the kernels and the C ABI");
// the deep point pass's buffers, sized for `cells` cells
static int deepBuffers(Sim* s, uint32_t cells) {
  if (s->deepCells >= cells && s->deepOut) return ST_OK;
  HIPCHK(hipStreamSynchronize(s->st));
  dfree(s->deepFlag); dfree(s->deepOff); dfree(s->deepSums); dfree(s->deepAll); dfree(s->deepOut);
  s->deepCells = 0;
  int rc;
  if ((rc = dalloc(&s->deepFlag, cells)) || (rc = dalloc(&s->deepOff, size_t(cells) + 1)) ||
      (rc = dalloc(&s->deepSums, size_t(cells) / 1024 + 2)) || (rc = dalloc(&s->deepAll, 1)) ||
      (rc = dalloc(&s->deepOut, size_t(ST_DEEP_MAX))))
    return rc;
  s->deepCells = cells;
  return ST_OK;
}
int st_read_deep(void* p, float* out4, uint32_t room, uint32_t* n) {
  Sim* s = getSim(p);
  if (!s || !n || (room && !out4)) return fail(ST_ERR_ARG, "st_read_deep: null argument");
  *n = 0;
  if (!s->nbOn || !s->G.velG || !s->N || !s->cur) return ST_OK;  // readDeep gives null (sim.js:409)
  HIPCHK(hipSetDevice(s->dev));
  const uint32_t cells = s->scap * 512u;
  int rc;
  if ((rc = deepBuffers(s, cells)) != ST_OK) return rc;
  launch_deep(s->st, s->ctx(), cells, s->deepPrev->load(), s->deepFlag, s->deepOff, s->deepSums, s->deepAll, s->deepOut);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(ST_ERR_HIP, std::string("st_read_deep: ") + hipGetErrorString(e));
  uint32_t kept = 0, all = 0;
  HIPCHK(hipMemcpyAsync(&kept, s->deepOff + cells, 4, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipMemcpyAsync(&all, s->deepAll, 4, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  s->deepPrev->store(all);  // deepFin: this call's count picks the next call's sample
  const uint32_t m = std::min<uint32_t>(std::min<uint32_t>(kept, ST_DEEP_MAX), room);
  if (m) HIPCHK(hipMemcpy(out4, s->deepOut, size_t(m) * 16, hipMemcpyDeviceToHost));
  *n = m;
  return ST_OK;
}

// ---------------------------------------------------------------------------
// asynchronous readback (st_read_begin / st_read_wait / st_read_free)

struct ReadTicket {
  int dev = 0;
  uint32_t what = 0, n = 0;
  hipEvent_t ev = nullptr;
  char* buf = nullptr;
  size_t bytes = 0;
  std::shared_ptr<PinCache> cache;  // set when buf is the cache's
  std::shared_ptr<std::atomic<uint32_t>> deepPrev;
  ~ReadTicket() {
    if (ev) {
      (void)hipEventSynchronize(ev);  // the copy into buf has landed before buf is let go
      (void)hipEventDestroy(ev);
    }
    if (buf && cache) {
      std::lock_guard<std::mutex> g(cache->m);
      cache->lent = false;
    } else if (buf) {
      (void)hipHostFree(buf);
    }
  }
};

static int ticketBuffer(Sim* s, ReadTicket* t, size_t bytes) {
  bytes = std::max<size_t>(bytes, 64);
  if (!s->pin) s->pin = std::make_shared<PinCache>();
  PinCache& c = *s->pin;
  std::lock_guard<std::mutex> g(c.m);
  if (!c.lent) {
    if (c.bytes < bytes) {
      if (c.ptr) (void)hipHostFree(c.ptr);
      c.ptr = nullptr;
      c.bytes = 0;
      if (hipHostMalloc((void**)&c.ptr, bytes, hipHostMallocDefault) != hipSuccess) {
        c.ptr = nullptr;
        return fail(ST_ERR_NOMEM, "st_read_begin: hipHostMalloc failed");
      }
      c.bytes = bytes;
    }
    c.lent = true;
    t->buf = c.ptr;
    t->cache = s->pin;
  } else if (hipHostMalloc((void**)&t->buf, bytes, hipHostMallocDefault) != hipSuccess) {
    t->buf = nullptr;
    return fail(ST_ERR_NOMEM, "st_read_begin: hipHostMalloc failed");
  }
  t->bytes = bytes;
  return ST_OK;
}

int st_read_begin(void* p, uint32_t what, uint32_t n, void** ticket) {
  Sim* s = getSim(p);
  if (!s || !ticket) return fail(ST_ERR_ARG, "st_read_begin: null argument");
  *ticket = nullptr;
  if (what > ST_READ_DEEP) return fail(ST_ERR_ARG, "st_read_begin: unknown result");
  HIPCHK(hipSetDevice(s->dev));
  std::unique_ptr<ReadTicket> t(new ReadTicket());
  t->dev = s->dev;
  t->what = what;
  int rc;
  if (what == ST_READ_PARTICLES) {
    t->n = std::min(n, s->N);
    if ((rc = ticketBuffer(s, t.get(), size_t(t->n) * 16)) != ST_OK) return rc;
    if (t->n) {
      launch_unsort(s->st, s->G.pos, s->G.pid, s->G.scratch4, t->n);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(t->buf, s->G.scratch4, size_t(t->n) * 16, hipMemcpyDeviceToHost, s->st));
    }
  } else if (what == ST_READ_CTL) {
    if ((rc = ticketBuffer(s, t.get(), 64)) != ST_OK) return rc;
    HIPCHK(hipMemcpyAsync(t->buf, s->G.ctl, 64, hipMemcpyDeviceToHost, s->st));
  } else if (what == ST_READ_SPRAY) {
    // the slot counter, then every slot (positions, velocities): n is fixed at the wait
    t->n = s->sprayCap;
    if ((rc = ticketBuffer(s, t.get(), 16 + size_t(s->sprayCap) * 32)) != ST_OK) return rc;
    if (s->sprayCap) {
      HIPCHK(hipMemcpyAsync(t->buf, s->G.sctl, 4, hipMemcpyDeviceToHost, s->st));
      HIPCHK(hipMemcpyAsync(t->buf + 16, s->G.sPos, size_t(s->sprayCap) * 16, hipMemcpyDeviceToHost, s->st));
      HIPCHK(hipMemcpyAsync(t->buf + 16 + size_t(s->sprayCap) * 16, s->G.sVel, size_t(s->sprayCap) * 16, hipMemcpyDeviceToHost, s->st));
    }
  } else {  // ST_READ_DEEP: the kept and counted cells, then the points
    const bool none = !s->nbOn || !s->G.velG || !s->N || !s->cur;  // readDeep gives null (sim.js:409)
    t->n = none ? 0u : 1u;
    if ((rc = ticketBuffer(s, t.get(), none ? 64 : 16 + size_t(ST_DEEP_MAX) * 16)) != ST_OK) return rc;
    if (!none) {
      const uint32_t cells = s->scap * 512u;
      if ((rc = deepBuffers(s, cells)) != ST_OK) return rc;
      launch_deep(s->st, s->ctx(), cells, s->deepPrev->load(), s->deepFlag, s->deepOff, s->deepSums, s->deepAll, s->deepOut);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(t->buf, s->deepOff + cells, 4, hipMemcpyDeviceToHost, s->st));
      HIPCHK(hipMemcpyAsync(t->buf + 4, s->deepAll, 4, hipMemcpyDeviceToHost, s->st));
      HIPCHK(hipMemcpyAsync(t->buf + 16, s->deepOut, size_t(ST_DEEP_MAX) * 16, hipMemcpyDeviceToHost, s->st));
      t->deepPrev = s->deepPrev;
    }
  }
  HIPCHK(hipEventCreateWithFlags(&t->ev, hipEventDisableTiming));
  HIPCHK(hipEventRecord(t->ev, s->st));
  *ticket = t.release();
  return ST_OK;
}

int st_read_wait(void* p, void* out0, void* out1, uint32_t room, uint32_t* n) {
  ReadTicket* t = (ReadTicket*)p;
  if (!t || !n) return fail(ST_ERR_ARG, "st_read_wait: null argument");
  *n = 0;
  HIPCHK(hipSetDevice(t->dev));
  HIPCHK(hipEventSynchronize(t->ev));
  switch (t->what) {
    case ST_READ_PARTICLES:
      if (t->n > room || (t->n && !out0)) return fail(ST_ERR_ARG, "st_read_wait: output array too small");
      if (t->n) std::memcpy(out0, t->buf, size_t(t->n) * 16);
      *n = t->n;
      break;
    case ST_READ_CTL:
      if (room < 16 || !out0) return fail(ST_ERR_ARG, "st_read_wait: the control words need 16 entries");
      std::memcpy(out0, t->buf, 64);
      *n = 16;
      break;
    case ST_READ_SPRAY: {
      if (!t->n) break;
      uint32_t used = 0;
      std::memcpy(&used, t->buf, 4);
      const uint32_t k = std::max(1u, std::min(t->n, used));
      if (k > room || !out0 || !out1) return fail(ST_ERR_ARG, "st_read_wait: output arrays too small");
      std::memcpy(out0, t->buf + 16, size_t(k) * 16);
      std::memcpy(out1, t->buf + 16 + size_t(t->n) * 16, size_t(k) * 16);
      *n = k;
      break;
    }
    default: {  // ST_READ_DEEP
      if (!t->n) break;
      uint32_t kept = 0, all = 0;
      std::memcpy(&kept, t->buf, 4);
      std::memcpy(&all, t->buf + 4, 4);
      t->deepPrev->store(all);  // deepFin: this read's count picks the next read's sample
      const uint32_t m = std::min<uint32_t>(std::min<uint32_t>(kept, ST_DEEP_MAX), room);
      if (m && !out0) return fail(ST_ERR_ARG, "st_read_wait: null output");
      if (m) std::memcpy(out0, t->buf + 16, size_t(m) * 16);
      *n = m;
    }
  }
  return ST_OK;
}

int st_read_free(void* p) {
  delete (ReadTicket*)p;
  return ST_OK;
}

int st_pack_view(void* p, uint32_t disp_max, uint32_t* out, size_t cap_words, st_view_layout* Lo) {
  Sim* s = getSim(p);
  if (!s || !Lo) return fail(ST_ERR_ARG, "st_pack_view: null argument");
  const uint32_t dispMax = disp_max ? disp_max : 300000u, N = s->N;
  const uint32_t stride = std::max<uint32_t>(1u, uint32_t((uint64_t(N) + dispMax - 1) / dispMax));
  const uint32_t sdMax = std::min<uint32_t>(s->sprayCap, 200000u);
  const bool spr = sdMax > 0 && N > 0, fromPos = !spr && (stride <= 1 || !N);
  const uint32_t nW = fromPos ? std::min<uint32_t>(N, std::min<uint32_t>(dispMax + 1, s->maxn)) : uint32_t((uint64_t(N) + stride - 1) / stride);
  const uint32_t nS = spr ? sdMax : 0u;
  // packLayout (sim.js:376-380)
  const uint32_t gW = (nW + 3u) / 4u, gS = (nS + 3u) / 4u, gD = 0u;
  st_view_layout L{};
  L.nW = nW; L.nS = nS; L.nD = 0;
  L.oWp = 4; L.oWs = L.oWp + gW * 6u; L.oWf = L.oWs + gW; L.oSp = L.oWf + gW; L.oSk = L.oSp + gS * 6u;
  L.oDp = L.oSk + gS * 4u; L.oDs = L.oDp + gD * 6u; L.words = L.oDs + gD;
  L.stride = fromPos ? 1u : stride;
  L.from_pos = fromPos ? 1u : 0u;
  *Lo = L;
  if (!out) return ST_OK;
  if (cap_words < L.words) return fail(ST_ERR_ARG, "st_pack_view: output shorter than layout.words");
  if (spr && !s->cur) return fail(ST_ERR_STATE, "st_pack_view: the whitewater sample needs a substep of this sim first");
  HIPCHK(hipSetDevice(s->dev));
  int rc;
  if (s->viewWords < L.words) {
    HIPCHK(hipStreamSynchronize(s->st));
    dfree(s->viewOut);
    s->viewWords = 0;
    if ((rc = dalloc(&s->viewOut, L.words))) return rc;
    s->viewWords = L.words;
  }
  if (!s->viewOff || s->viewSd < sdMax) {
    HIPCHK(hipStreamSynchronize(s->st));
    dfree(s->viewFlag); dfree(s->viewOff); dfree(s->viewSums); dfree(s->viewDisp);
    s->viewSd = 0;
    if ((rc = dalloc(&s->viewFlag, std::max(sdMax, 1u))) || (rc = dalloc(&s->viewOff, size_t(sdMax) + 1)) ||
        (rc = dalloc(&s->viewSums, size_t(sdMax) / 1024 + 2)) || (rc = dalloc(&s->viewDisp, std::max(sdMax, 1u))))
      return rc;
    s->viewSd = sdMax;
  }
  ViewParams V{};
  V.x0 = s->x0; V.y0 = s->y0; V.z0 = s->z0;
  V.kx = float(65535.0 / (double(s->x1) - double(s->x0)));
  V.ky = float(65535.0 / (double(s->y1) - double(s->y0)));
  V.kz = float(65535.0 / (double(s->z1) - double(s->z0)));
  V.vs = s->hp.vs != 0.0f ? s->hp.vs : 1.0f;
  V.foamLife = s->hp.wwFoamLife;
  V.nW = nW; V.nS = nS; V.srcStride = L.stride; V.fromPos = L.from_pos;
  V.oWp = L.oWp; V.oWs = L.oWs; V.oWf = L.oWf; V.oSp = L.oSp; V.oSk = L.oSk;
  V.n = N;
  LaunchCtx C = s->ctx();
  launch_pack_view(s->st, C, V, spr ? sdMax : 0u, s->viewFlag, s->viewOff, s->viewSums, s->viewDisp, s->viewOut);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, s->viewOut, size_t(L.words) * 4, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return ST_OK;
}

// the norms kernels with partials in `part` and results in `res`, read back into *out
static int normsInto(Sim* s, const LaunchCtx& L, double* part, double* res, st_norms* out) {
  launch_norms(s->st, L, part, res);
  HIPCHK(hipGetLastError());
  double h[ST_NORM_WORDS];
  HIPCHK(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  out->res_l2 = std::sqrt(h[0]);
  out->res_max = h[1];
  out->rhs_l2 = std::sqrt(h[2]);
  out->div_l2 = std::sqrt(h[3]);
  out->div_max = h[4];
  out->err_l2 = std::sqrt(h[5]);
  out->err_max = h[6];
  out->fluid = h[7];
  return ST_OK;
}

int st_solve_norms(void* p, st_norms* out) {
  Sim* s = getSim(p);
  if (!s || !out) return fail(ST_ERR_ARG, "st_solve_norms: null argument");
  if (!s->cur) return fail(ST_ERR_STATE, "st_solve_norms: no substep of this sim to measure (none yet, or the last one was batched)");
  HIPCHK(hipSetDevice(s->dev));
  LaunchCtx L = s->ctx();
  if (s->normCap < L.cellGrid) {
    HIPCHK(hipStreamSynchronize(s->st));
    dfree(s->normPart);
    s->normCap = 0;
    int rc = dalloc(&s->normPart, size_t(L.cellGrid + 1) * ST_NORM_WORDS);
    if (rc) return rc;
    s->normCap = L.cellGrid;
  }
  return normsInto(s, L, s->normPart, s->normPart + size_t(s->normCap) * ST_NORM_WORDS, out);
}

int st_solve_norms_static(void* p, st_norms* out) {
  Sim* s = getSim(p);
  if (!s || !out) return fail(ST_ERR_ARG, "st_solve_norms_static: null argument");
  if (!s->cur) return fail(ST_ERR_STATE, "st_solve_norms_static: no substep of this sim to measure (none yet, or the last one was batched)");
  HIPCHK(hipSetDevice(s->dev));
  LaunchCtx L = s->ctx();
  double* part = nullptr;
  size_t words = 0;
  HIPCHK(norms_static_scratch(s->dev, &part, &words));
  const size_t need = size_t(L.cellGrid + 1) * ST_NORM_WORDS;
  if (need > words)
    return fail(ST_ERR_ARG, "st_solve_norms_static: the launch needs " + std::to_string(need) + " scratch words, the fixed array holds " +
                                std::to_string(words));
  return normsInto(s, L, part, part + size_t(L.cellGrid) * ST_NORM_WORDS, out);
}

int st_sor_path(void* p, uint32_t* path) {
  Sim* s = getSim(p);
  if (!s || !path) return fail(ST_ERR_ARG, "st_sor_path: null argument");
  HIPCHK(hipSetDevice(s->dev));
  HIPCHK(hipStreamSynchronize(s->st));
  HIPCHK(hipMemcpy(path, s->G.sorBar + COOP_MAX_WG + 3, 4, hipMemcpyDeviceToHost));
  return ST_OK;
}

int st_mem_bytes(void* p, uint64_t* bytes) {
  Sim* s = getSim(p);
  if (!s || !bytes) return fail(ST_ERR_ARG, "st_mem_bytes: null argument");
  uint64_t M = s->maxn, C = uint64_t(s->scap) * 512;
  *bytes = M * (2 * (16 + 16 + 4) + 4 + 16) + uint64_t(s->ntiles) * 24 + C * (4 * 5 + 24 + (s->velGOn ? 12 : 0)) +
           uint64_t(s->scap) * (36 + P2G_HALO_WORDS * 4) + s->sdfCap * 4;
  return ST_OK;
}

int st_batch_create(void* const* sims, uint32_t n, void** out) {
  if (!sims || !out || n == 0) return fail(ST_ERR_ARG, "st_batch_create: null argument or no tanks");
  *out = nullptr;
  if (n > ST_BATCH_MAX) return fail(ST_ERR_ARG, "st_batch_create: at most 256 tanks");
  for (uint32_t k = 0; k < n; k++) {
    Sim* s = getSim(sims[k]);
    if (!s) return fail(ST_ERR_ARG, "st_batch_create: null tank");
    if (s->batch) return fail(ST_ERR_STATE, "st_batch_create: a tank is already in a batch");
    if (s->dev != getSim(sims[0])->dev) return fail(ST_ERR_ARG, "st_batch_create: tanks on different devices");
    for (uint32_t q = 0; q < k; q++)
      if (sims[q] == sims[k]) return fail(ST_ERR_ARG, "st_batch_create: a tank is listed twice");
  }
  Batch* b = new Batch();
  for (int k = 0; k < BRING; k++) b->evUsed[k] = false;
  b->dev = getSim(sims[0])->dev;
  hipError_t e = hipSetDevice(b->dev);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete b;
    return fail(ST_ERR_HIP, std::string("st_batch_create: ") + hipGetErrorString(e));
  }
  for (uint32_t k = 0; k < n; k++) {
    Sim* s = getSim(sims[k]);
    e = hipStreamSynchronize(s->st);
    if (e != hipSuccess) {
      for (uint32_t q = 0; q < k; q++) {
        Sim* r = b->tanks[q];
        r->st = r->ownStream;
        r->ownStream = nullptr;
        r->batch = nullptr;
      }
      (void)hipStreamDestroy(b->st);
      delete b;
      return fail(ST_ERR_HIP, std::string("st_batch_create: ") + hipGetErrorString(e));
    }
    s->ownStream = s->st;
    s->st = b->st;
    s->batch = b;
    b->tanks.push_back(s);
  }
  b->sorResident = b->tanks[0]->sorResident;
  b->coopGrid = b->tanks[0]->coopGrid;
  b->plan.resize(n);
  if (b->coopGrid && dalloc(&b->dPlan, size_t(n) * 3 + 1) != ST_OK) {
    (void)hipGetLastError();
    b->coopGrid = 0;  // launches per iteration
  }
  *out = b;
  return ST_OK;
}

int st_batch_destroy(void* p) {
  Batch* b = (Batch*)p;
  if (!b) return fail(ST_ERR_ARG, "st_batch_destroy: null batch");
  (void)hipSetDevice(b->dev);
  (void)hipStreamSynchronize(b->st);
  for (Sim* s : b->tanks) {
    s->st = s->ownStream;
    s->ownStream = nullptr;
    s->batch = nullptr;
  }
  freeRing(b);
  dfree(b->dPlan);
  (void)hipStreamDestroy(b->st);
  delete b;
  return ST_OK;
}

static int batchStep(Batch* b, const st_params* prm, uint32_t nprm, bool graph) {
  const uint32_t nt = uint32_t(b->tanks.size());
  if (nprm != 1u && nprm != nt) return fail(ST_ERR_ARG, "st_batch_step: give one parameter block or one per tank");
  HIPCHK(hipSetDevice(b->dev));
  uint32_t itersMax = 0;
  for (uint32_t t = 0; t < nt; t++) {
    const st_params* o = prm + (nprm == 1u ? 0u : t);
    if (o->nb_on || o->ocean_on || o->only_tiles)
      return fail(ST_ERR_ARG, "st_batch_step: the narrow band, the ocean and tile-only steps are single-tank only");
    if (o->extrap_layers) return fail(ST_ERR_ARG, "st_batch_step: velocity extrapolation is single-tank only");
    if (b->tanks[t]->sprayCap) return fail(ST_ERR_ARG, "st_batch_step: whitewater is single-tank only (a tank has a spray pool)");
    itersMax = std::max<uint32_t>(itersMax, b->tanks[t]->N ? uint32_t(std::max(1, int(o->iters))) : 0u);
  }
  int rc = ensureRing(b, itersMax);
  if (rc) return rc;
  if (graph && !b->dGraph && (rc = dalloc(&b->dGraph, b->slotBytes))) return rc;
  uint32_t k = b->ringPos % BRING;
  if (b->evUsed[k]) HIPCHK(hipEventSynchronize(b->ev[k]));
  uint8_t* h = b->hRing + k * b->slotBytes;
  uint8_t* d = graph ? b->dGraph : b->dRing + k * b->slotBytes;
  BTank* hT = (BTank*)h;
  DParams* hP = (DParams*)(h + b->offP);
  DCols* hC = (DCols*)(h + b->offC);
  uint32_t* hTab = (uint32_t*)(h + b->offTab);
  for (uint32_t t = 0; t < nt; t++) {
    Sim* s = b->tanks[t];
    BTankPlan& pl = b->plan[t];
    pl = BTankPlan{0u, s->ntiles, s->scap, 0u, false, false};
    hT[t].P = (const DParams*)(d + b->offP) + t;
    hT[t].C = (const DCols*)(d + b->offC) + t;
    hT[t].G = s->G;
    if (!s->N) continue;  // step() does nothing when N = 0
    if ((rc = prepStep(s, prm + (nprm == 1u ? 0u : t), hP[t], hC[t]))) return rc;
    pl.n = s->N;
    pl.iters = uint32_t(s->iters);
    pl.binned = s->binned;
    pl.solid = s->solidDirty;
  }
  batch_tables(b->plan.data(), nt, itersMax, b->sorResident, hTab);
  HIPCHK(hipMemcpyAsync(d, h, b->offTab + batch_table_words(nt, itersMax) * 4u, hipMemcpyHostToDevice, b->st));
  if (!b->evUsed[k]) {
    HIPCHK(hipEventCreateWithFlags(&b->ev[k], hipEventDisableTiming));
    b->evUsed[k] = true;
  }
  HIPCHK(hipEventRecord(b->ev[k], b->st));
  b->ringPos++;
  if (!graph) {
    b->lastResident = launch_batch_step(b->st, (const BTank*)d, (const uint32_t*)(d + b->offTab), hTab, nt, itersMax, b->coopGrid,
                                        b->dPlan) ? 1u : 0u;
  } else {
    // the block tables fix every launch's grid and kernel choice; the tank
    // records, parameters and tables are read from the graph slot at run time.
    // The SOR takes one launch per iteration in a graph: a captured
    // cooperative launch is not scheduled as one when other streams run.
    b->lastResident = 0u;
    const size_t words = batch_table_words(nt, itersMax);
    Batch::Captured* hit = nullptr;
    for (auto& g : b->graphs)
      if (g.tab.size() == words && std::memcmp(g.tab.data(), hTab, words * 4u) == 0) hit = &g;
    if (hit) {
      HIPCHK(hipGraphLaunch(hit->exec, b->st));
    } else {
      HIPCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
      launch_batch_step(b->st, (const BTank*)d, (const uint32_t*)(d + b->offTab), hTab, nt, itersMax, 0u, nullptr);
      hipError_t le = hipGetLastError();
      hipGraph_t g = nullptr;
      hipError_t ce = hipStreamEndCapture(b->st, &g);
      if (le != hipSuccess || ce != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(ST_ERR_HIP, std::string("st_batch_step_n: capture: ") + hipGetErrorString(le != hipSuccess ? le : ce));
      }
      Batch::Captured c;
      c.tab.assign(hTab, hTab + words);
      hipError_t ie = hipGraphInstantiate(&c.exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) return fail(ST_ERR_HIP, std::string("st_batch_step_n: hipGraphInstantiate: ") + hipGetErrorString(ie));
      HIPCHK(hipGraphLaunch(c.exec, b->st));
      if (b->graphs.size() >= GRAPH_CACHE) {
        HIPCHK(hipStreamSynchronize(b->st));
        (void)hipGraphExecDestroy(b->graphs.front().exec);
        b->graphs.erase(b->graphs.begin());
      }
      b->graphs.push_back(std::move(c));
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(ST_ERR_HIP, std::string("batched launch: ") + hipGetErrorString(e));
  // what the single-tank stages leave behind: g2p wrote the new state into the
  // other buffers and binned it for the next substep
  for (uint32_t t = 0; t < nt; t++) {
    Sim* s = b->tanks[t];
    if (!s->N) continue;
    s->solidDirty = false;
    s->sorSeq++;
    s->binned = true;
    s->cur = nullptr;
    s->begun = false;
    std::swap(s->G.pos, s->G.posB);
    std::swap(s->G.vel, s->G.velB);
    std::swap(s->G.pid, s->G.pidB);
  }
  return ST_OK;
}

int st_batch_step(void* p, const st_params* prm, uint32_t nprm) {
  Batch* b = (Batch*)p;
  if (!b || !prm) return fail(ST_ERR_ARG, "st_batch_step: null argument");
  return batchStep(b, prm, nprm, false);
}

int st_batch_step_n(void* p, const st_params* prm, uint32_t nprm, int32_t nsteps, int32_t use_graph) {
  Batch* b = (Batch*)p;
  if (!b || !prm) return fail(ST_ERR_ARG, "st_batch_step_n: null argument");
  for (int32_t k = 0; k < nsteps; k++) {
    int rc = batchStep(b, prm, nprm, use_graph != 0);
    if (rc) return rc;
  }
  return ST_OK;
}

int st_batch_sync(void* p) {
  Batch* b = (Batch*)p;
  if (!b) return fail(ST_ERR_ARG, "st_batch_sync: null batch");
  HIPCHK(hipSetDevice(b->dev));
  HIPCHK(hipStreamSynchronize(b->st));
  return ST_OK;
}

int st_batch_sor_grid(void* p, uint32_t workgroups) {
  Batch* b = (Batch*)p;
  if (!b) return fail(ST_ERR_ARG, "st_batch_sor_grid: null batch");
  const uint32_t most = b->tanks[0]->coopGrid;
  if (workgroups && !b->dPlan) {
    if (!most) return fail(ST_ERR_STATE, "st_batch_sor_grid: cooperative launches are not available on this device");
    HIPCHK(hipSetDevice(b->dev));
    int rc = dalloc(&b->dPlan, b->tanks.size() * 3 + 1);
    if (rc) return rc;
  }
  // a tank that streams tile windows runs four pipelines per workgroup over
  // eight XCD groups, which needs at least two workgroups
  b->coopGrid = workgroups ? std::max(2u, std::min(workgroups, most)) : 0u;
  return ST_OK;
}

int st_batch_sor_path(void* p, uint32_t* resident) {
  Batch* b = (Batch*)p;
  if (!b || !resident) return fail(ST_ERR_ARG, "st_batch_sor_path: null argument");
  *resident = b->lastResident;
  return ST_OK;
}

int st_batch_size(void* p, uint32_t* n) {
  Batch* b = (Batch*)p;
  if (!b || !n) return fail(ST_ERR_ARG, "st_batch_size: null argument");
  *n = uint32_t(b->tanks.size());
  return ST_OK;
}

int st_snapshot_create(void* p, void** out) {
  Sim* s = getSim(p);
  if (!s || !out) return fail(ST_ERR_ARG, "st_snapshot_create: null argument");
  *out = nullptr;
  HIPCHK(hipSetDevice(s->dev));
  Snap* q = new Snap();
  size_t M = s->maxn, C = size_t(s->scap) * 512u;
  int rc;
  if ((rc = dalloc(&q->pos, M)) || (rc = dalloc(&q->vel, M)) || (rc = dalloc(&q->pid, M)) || (rc = dalloc(&q->tmap, s->ntiles)) ||
      (rc = dalloc(&q->slotTile, s->scap)) || (rc = dalloc(&q->fstack, s->scap)) || (rc = dalloc(&q->slotList, s->scap)) ||
      (rc = dalloc(&q->ctl, 16)) || (rc = dalloc(&q->ctype, C)) || (rc = dalloc(&q->pres, C))) {
    freeSnap(q);
    return rc;
  }
  D2D(q->pos, s->G.pos, M); D2D(q->vel, s->G.vel, M); D2D(q->pid, s->G.pid, M); D2D(q->tmap, s->G.tmap, s->ntiles);
  D2D(q->slotTile, s->G.slotTile, s->scap); D2D(q->fstack, s->G.fstack, s->scap); D2D(q->slotList, s->G.slotList, s->scap);
  D2D(q->ctl, s->G.ctl, 16); D2D(q->ctype, s->G.ctype, C); D2D(q->pres, s->G.pres, C);
  HIPCHK(hipStreamSynchronize(s->st));
  q->N = s->N; q->scap = s->scap; q->seed = s->seed; q->tm = s->tm;
  *out = q;
  return ST_OK;
}

int st_snapshot_restore(void* p, void* snap) {
  Sim* s = getSim(p);
  Snap* q = (Snap*)snap;
  if (!s || !q) return fail(ST_ERR_ARG, "st_snapshot_restore: null argument");
  if (q->scap != s->scap) return fail(ST_ERR_STATE, "st_snapshot_restore: the pool has grown since the snapshot");
  HIPCHK(hipSetDevice(s->dev));
  size_t M = s->maxn, C = size_t(s->scap) * 512u;
  D2D(s->G.pos, q->pos, M); D2D(s->G.vel, q->vel, M); D2D(s->G.pid, q->pid, M); D2D(s->G.tmap, q->tmap, s->ntiles);
  D2D(s->G.slotTile, q->slotTile, s->scap); D2D(s->G.fstack, q->fstack, s->scap); D2D(s->G.slotList, q->slotList, s->scap);
  D2D(s->G.ctl, q->ctl, 16); D2D(s->G.ctype, q->ctype, C); D2D(s->G.pres, q->pres, C);
  s->N = q->N; s->seed = q->seed; s->tm = q->tm;
  s->solidDirty = true;  // the restored cell types hold the collider bits of the records then
  s->binned = false;  // marks and bins are rebuilt from the restored positions
  s->begun = false;
  return ST_OK;
}

int st_snapshot_destroy(void* snap) {
  if (!snap) return fail(ST_ERR_ARG, "st_snapshot_destroy: null snapshot");
  freeSnap((Snap*)snap);
  return ST_OK;
}

int st_tune_sor(void* p, const st_params* prm, int32_t reps, st_sor_tune* out) {
  Sim* s = getSim(p);
  if (!s || !prm || !out) return fail(ST_ERR_ARG, "st_tune_sor: null argument");
  if (s->batch) return fail(ST_ERR_STATE, "st_tune_sor: the tank is in a batch (use st_batch_tune_sor)");
  if (!s->N) return fail(ST_ERR_STATE, "st_tune_sor: no particles");
  reps = std::max(1, std::min(reps, 50));
  HIPCHK(hipSetDevice(s->dev));
  std::memset(out, 0, sizeof *out);
  // candidates: the two cooperative paths, then one launch per iteration at
  // 1, 1/2 and 1/4 of the resident width (whole groups of 8, one per XCD)
  uint32_t def = std::min<uint32_t>((s->scap + 7u) & ~7u, s->sorResident);
  uint32_t n = 0;
  if (s->coopGrid) {
    out->cand_mode[n] = 0; out->cand_width[n++] = 0;
    out->cand_mode[n] = 2; out->cand_width[n++] = 0;
  }
  for (uint32_t div = 1; div <= 4; div *= 2) {
    uint32_t w = std::max<uint32_t>(8u, (def / div) & ~7u);
    bool dup = false;
    for (uint32_t q = 0; q < n; q++) dup |= out->cand_mode[q] == 1u && out->cand_width[q] == w;
    if (!dup) { out->cand_mode[n] = 1; out->cand_width[n++] = w; }
  }
  out->n = n;
  void* snap = nullptr;
  int rc = st_snapshot_create(p, &snap);
  if (rc) return rc;
  st_params q = *prm;
  q.nb_on = 0; q.nb_init = 0; q.only_tiles = 0; q.sor_mode = 0; q.ocean_on = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  rc = st_begin_step(p, &q);
  for (int k = 0; !rc && k <= ST_STAGE_DIVERGENCE; k++) rc = runStage(s, k);
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = fail(ST_ERR_HIP, "st_tune_sor: hipEventCreate");
  for (uint32_t c = 0; !rc && c < n; c++) {
    s->sorMode = out->cand_mode[c];
    s->pipeOverride = out->cand_width[c];
    rc = runStage(s, ST_STAGE_PRESSURE);  // warm-up
    if (!rc && hipEventRecord(e0, s->st) != hipSuccess) rc = fail(ST_ERR_HIP, "st_tune_sor: hipEventRecord");
    for (int32_t r = 0; !rc && r < reps; r++) rc = runStage(s, ST_STAGE_PRESSURE);
    if (!rc && (hipEventRecord(e1, s->st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)) rc = fail(ST_ERR_HIP, "st_tune_sor: timing");
    float ms = 0.0f;
    if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = fail(ST_ERR_HIP, "st_tune_sor: hipEventElapsedTime");
    out->cand_ms[c] = ms / float(reps);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  // back to the state the tuning started from; the collider lookup is rebuilt
  // on the next substep (the tuning step may have taken the records as seen)
  s->pipeOverride = 0;
  int rc2 = st_snapshot_restore(p, snap);
  (void)hipStreamSynchronize(s->st);  // the restore copies read the snapshot
  (void)st_snapshot_destroy(snap);
  s->solidDirty = true;
  s->cur = nullptr;
  if (!rc) rc = rc2;
  if (!rc) rc = st_sync(p);
  if (rc) return rc;
  uint32_t best = 0;
  for (uint32_t c = 1; c < n; c++)
    if (out->cand_ms[c] < out->cand_ms[best]) best = c;
  out->mode = out->cand_mode[best];
  out->width = out->cand_width[best];
  s->tuned = true;
  s->tunedMode = out->mode;
  s->tunedPipe = out->width;
  return ST_OK;
}

int st_batch_tune_sor(void* p, const st_params* prm, uint32_t nprm, int32_t reps, st_sor_tune* out) {
  Batch* b = (Batch*)p;
  if (!b || !prm || !out) return fail(ST_ERR_ARG, "st_batch_tune_sor: null argument");
  reps = std::max(1, std::min(reps, 50));
  HIPCHK(hipSetDevice(b->dev));
  std::memset(out, 0, sizeof *out);
  const uint32_t nt = uint32_t(b->tanks.size());
  uint32_t base = b->tanks[0]->sorResident, n = 0;
  const uint32_t coop = b->dPlan ? b->tanks[0]->coopGrid : 0u, keepCoop = b->coopGrid;
  if (coop) {  // the resident solves in one cooperative launch
    out->cand_mode[n] = 0;
    out->cand_width[n++] = coop;
  }
  for (uint32_t mul = 1; mul <= 8; mul *= 2) {  // one launch per iteration, 1/4, 1/2, 1, 2 of the resident width
    out->cand_mode[n] = 1;
    out->cand_width[n++] = std::max<uint32_t>(8u, (base * mul / 4u) & ~7u);
  }
  out->n = n;
  std::vector<void*> snaps(nt, nullptr);
  int rc = ST_OK;
  for (uint32_t t = 0; !rc && t < nt; t++) rc = st_snapshot_create(b->tanks[t], &snaps[t]);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = fail(ST_ERR_HIP, "st_batch_tune_sor: hipEventCreate");
  uint32_t keep = b->sorResident;
  for (uint32_t c = 0; !rc && c < n; c++) {
    if (out->cand_mode[c] == 0) {
      b->coopGrid = out->cand_width[c];
    } else {
      b->coopGrid = 0u;
      b->sorResident = out->cand_width[c];
    }
    rc = st_batch_step(p, prm, nprm);  // warm-up
    if (!rc && hipEventRecord(e0, b->st) != hipSuccess) rc = fail(ST_ERR_HIP, "st_batch_tune_sor: hipEventRecord");
    for (int32_t r = 0; !rc && r < reps; r++) rc = st_batch_step(p, prm, nprm);
    if (!rc && (hipEventRecord(e1, b->st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)) rc = fail(ST_ERR_HIP, "st_batch_tune_sor: timing");
    float ms = 0.0f;
    if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = fail(ST_ERR_HIP, "st_batch_tune_sor: hipEventElapsedTime");
    out->cand_ms[c] = ms / float(reps);
    for (uint32_t t = 0; !rc && t < nt; t++) {
      rc = st_snapshot_restore(b->tanks[t], snaps[t]);
      b->tanks[t]->solidDirty = true;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (uint32_t t = 0; t < nt; t++)
    if (snaps[t]) {
      // the snapshot copies run on the batch stream: wait before freeing them
      (void)hipStreamSynchronize(b->st);
      (void)st_snapshot_destroy(snaps[t]);
    }
  if (rc) {
    b->sorResident = keep;
    b->coopGrid = keepCoop;
    return rc;
  }
  // the fastest overall for stream steps; the fastest per-iteration width for
  // graph replays (which always launch per iteration)
  uint32_t best = 0, bestPipe = n;
  for (uint32_t c = 1; c < n; c++)
    if (out->cand_ms[c] < out->cand_ms[best]) best = c;
  for (uint32_t c = 0; c < n; c++)
    if (out->cand_mode[c] == 1u && (bestPipe == n || out->cand_ms[c] < out->cand_ms[bestPipe])) bestPipe = c;
  out->mode = out->cand_mode[best];
  out->width = out->cand_width[best];
  b->sorResident = out->cand_width[bestPipe];
  b->coopGrid = out->mode == 0u ? out->width : 0u;
  return ST_OK;
}

int st_adj_create(void* p, uint32_t max_steps, void** out) {
  Sim* s = getSim(p);
  if (!s || !out || max_steps == 0) return fail(ST_ERR_ARG, "st_adj_create: null argument or no steps");
  *out = nullptr;
  HIPCHK(hipSetDevice(s->dev));
  Adj* a = new Adj();
  a->s = s;
  a->maxSteps = max_steps;
  a->scap = s->scap;
  size_t M = s->maxn, C = size_t(s->scap) * 512u;
  AdjGrad& G = a->G;
  int rc;
  if ((rc = dalloc(&G.gPos, M)) || (rc = dalloc(&G.gVel, M)) || (rc = dalloc(&G.gPos2, M)) || (rc = dalloc(&G.gVel2, M)) ||
      (rc = dalloc(&G.gFN, C * 3)) || (rc = dalloc(&G.gFO, C * 3)) || (rc = dalloc(&G.gPres, C)) || (rc = dalloc(&G.gPres2, C)) ||
      (rc = dalloc(&G.gRhs, C)) || (rc = dalloc(&G.carry, size_t(s->ntiles) * 512u)) || (rc = dalloc(&G.sorNb, C * 6)) ||
      (rc = dalloc(&G.sorCoef, C * 6)) || (rc = dalloc(&G.gCol, size_t(ADJ_MAX_COLS) * 9)) || (rc = dalloc(&G.bar, ADJ_BAR_WORDS))) {
    freeAdj(a);
    return rc;
  }
  {
    hipError_t e = hipMemsetAsync(G.bar, 0, ADJ_BAR_WORDS * sizeof(uint32_t), s->st);
    int cus = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->dev);
    if (e != hipSuccess) {
      freeAdj(a);
      return fail(ST_ERR_HIP, std::string("st_adj_create: ") + hipGetErrorString(e));
    }
    a->coopGrid = adj_sor_coop_grid(uint32_t(cus));
  }
  G.ncell = uint32_t(C);
  G.ncarry = s->ntiles * 512u;
  *out = a;
  return ST_OK;
}

// After a synchronisation: a time-out of the one-launch solve's grid barrier
// (a workgroup that never arrived) left gradients that are not valid.  It is
// reported once, and the context falls back to one launch per sweep.
static int adjBarrierCheck(Adj* a, const char* who) {
  uint32_t fails = 0;
  HIPCHK(hipMemcpy(&fails, a->G.bar + ADJ_BAR_FAILS, sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (!fails) return ST_OK;
  HIPCHK(hipMemset(a->G.bar + ADJ_BAR_FAILS, 0, sizeof(uint32_t)));
  a->coopGrid = 0;
  return fail(ST_ERR_HIP, std::string(who) + ": a grid barrier of the one-launch adjoint solve timed out " + std::to_string(fails) +
                              " times; the gradients are not valid (later solves use one launch per sweep)");
}

int st_adj_destroy(void* adj) {
  Adj* a = (Adj*)adj;
  if (!a) return fail(ST_ERR_ARG, "st_adj_destroy: null context");
  (void)hipStreamSynchronize(a->s->st);
  freeAdj(a);
  return ST_OK;
}

int st_adj_record(void* adj, const st_params* prm) {
  Adj* a = (Adj*)adj;
  if (!a || !prm) return fail(ST_ERR_ARG, "st_adj_record: null argument");
  Sim* s = a->s;
  if (a->depth >= a->maxSteps) return fail(ST_ERR_STATE, "st_adj_record: every tape is in use (reverse some substeps first)");
  if (s->scap != a->scap) return fail(ST_ERR_STATE, "st_adj_record: the pool has grown since the context was made");
  if (prm->nb_on || prm->ocean_on || prm->only_tiles || prm->drag != 0.0f || s->open != 0u)
    return fail(ST_ERR_ARG, "st_adj_record: the adjoint covers closed boxes without drag or the narrow band");
  if (prm->rk2 || prm->extrap_layers)
    return fail(ST_ERR_ARG, "st_adj_record: the adjoint follows the reference's Euler step without extrapolation");
  const uint32_t ncol = s->ncTotal >= prm->nfv ? s->ncTotal - prm->nfv : 0u;
  if (ncol > ADJ_MAX_COLS)
    return fail(ST_ERR_ARG, "st_adj_record: the adjoint follows at most " + std::to_string(ADJ_MAX_COLS) + " colliders");
  HIPCHK(hipSetDevice(s->dev));
  if (a->tapes.size() <= a->depth) {
    a->tapes.emplace_back();
    std::memset(&a->tapes.back(), 0, sizeof(AdjTape));
    a->sdfSig.push_back(0);
    a->sdfAt.push_back(nullptr);
    a->ncols.push_back(0);
    int rc = allocTape(s, a->tapes.back());
    if (rc) return rc;
  }
  AdjTape& T = a->tapes[a->depth];
  size_t M = s->maxn, C = size_t(s->scap) * 512u;
  D2D(T.pos, s->G.pos, M); D2D(T.vel, s->G.vel, M); D2D(T.pid, s->G.pid, M); D2D(T.tmap0, s->G.tmap, s->ntiles);
  int rc = st_begin_step(s, prm);
  if (rc) return rc;
  T.n = s->N;
  T.iters = uint32_t(s->iters);
  D2D(T.P, &s->cur->P, 1);
  D2D(T.C, &s->cur->C, 1);
  T.sdf = s->G.sdf;
  a->sdfSig[a->depth] = s->sdfSig;
  a->sdfAt[a->depth] = s->G.sdf;
  a->ncols[a->depth] = ncol;
  for (int k = 0; k < ST_STAGE_COUNT; k++) {
    if ((rc = runStage(s, k))) return rc;
    if (k == ST_STAGE_BLOCKS) {
      D2D(T.tmap, s->G.tmap, s->ntiles); D2D(T.slotTile, s->G.slotTile, s->scap); D2D(T.slotList, s->G.slotList, s->scap);
      D2D(T.binStart, s->G.binStart, size_t(s->ntiles) + 2); D2D(T.binIdx, s->G.binIdx, M);
    }
    if (k == ST_STAGE_P2G) D2D(T.accWraw, s->G.accW, C * 3);
    if (k == ST_STAGE_CLASSIFY) D2D(T.ctype, s->G.ctype, C);
    if (k == ST_STAGE_DIVERGENCE) D2D(T.sten, s->G.sten, C);
    if (k == ST_STAGE_PROJECT) { D2D(T.accVfin, s->G.accV, C * 3); D2D(T.accWfin, s->G.accW, C * 3); D2D(T.ctl, s->G.ctl, 16); }
  }
  a->depth++;
  return ST_OK;
}

int st_adj_seed(void* adj, const float* g_pos4, const float* g_vel4) {
  Adj* a = (Adj*)adj;
  if (!a || !g_pos4 || !g_vel4) return fail(ST_ERR_ARG, "st_adj_seed: null argument");
  Sim* s = a->s;
  HIPCHK(hipSetDevice(s->dev));
  HIPCHK(hipMemcpyAsync(a->G.gPos, g_pos4, size_t(s->N) * 16, hipMemcpyHostToDevice, s->st));
  HIPCHK(hipMemcpyAsync(a->G.gVel, g_vel4, size_t(s->N) * 16, hipMemcpyHostToDevice, s->st));
  HIPCHK(hipMemsetAsync(a->G.carry, 0, size_t(a->G.ncarry) * 4, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return ST_OK;
}

int st_adj_backward(void* adj) {
  Adj* a = (Adj*)adj;
  if (!a) return fail(ST_ERR_ARG, "st_adj_backward: null context");
  if (a->depth == 0) return fail(ST_ERR_STATE, "st_adj_backward: no recorded substep left to reverse");
  Sim* s = a->s;
  HIPCHK(hipSetDevice(s->dev));
  AdjTape& T = a->tapes[a->depth - 1];
  if (a->ncols[a->depth - 1] && (a->sdfSig[a->depth - 1] != s->sdfSig || a->sdfAt[a->depth - 1] != s->G.sdf))
    return fail(ST_ERR_STATE, "st_adj_backward: the collider grids changed since the substep was recorded");
  a->lastOne = launch_adjoint_step(s->st, T, a->G, a->scap, s->ntiles, a->oneLaunch ? a->coopGrid : 0u);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(ST_ERR_HIP, std::string("adjoint launch: ") + hipGetErrorString(e));
  std::swap(a->G.gPos, a->G.gPos2);
  std::swap(a->G.gVel, a->G.gVel2);
  a->lastCols = a->ncols[a->depth - 1];
  a->depth--;
  return ST_OK;
}

static_assert(ADJ_MAX_COLS == ST_ADJ_MAX_COLS, "one collider limit for the kernels and the C ABI");
int st_adj_col_grad(void* adj, float* g, uint32_t ncols) {
  Adj* a = (Adj*)adj;
  if (!a || (ncols && !g)) return fail(ST_ERR_ARG, "st_adj_col_grad: null argument");
  if (ncols != a->lastCols)
    return fail(ST_ERR_ARG, "st_adj_col_grad: the substep reversed last had " + std::to_string(a->lastCols) + " colliders");
  if (!ncols) return ST_OK;
  Sim* s = a->s;
  HIPCHK(hipSetDevice(s->dev));
  HIPCHK(hipMemcpyAsync(g, a->G.gCol, size_t(ncols) * 9 * sizeof(float), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return adjBarrierCheck(a, "st_adj_col_grad");
}

int st_adj_grad(void* adj, float* g_pos4, float* g_vel4) {
  Adj* a = (Adj*)adj;
  if (!a || !g_pos4 || !g_vel4) return fail(ST_ERR_ARG, "st_adj_grad: null argument");
  Sim* s = a->s;
  HIPCHK(hipSetDevice(s->dev));
  HIPCHK(hipMemcpyAsync(g_pos4, a->G.gPos, size_t(s->N) * 16, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipMemcpyAsync(g_vel4, a->G.gVel, size_t(s->N) * 16, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return adjBarrierCheck(a, "st_adj_grad");
}

int st_adj_set_sor_path(void* adj, uint32_t one_launch) {
  Adj* a = (Adj*)adj;
  if (!a) return fail(ST_ERR_ARG, "st_adj_set_sor_path: null context");
  a->oneLaunch = one_launch != 0u;
  return ST_OK;
}

int st_adj_sor_path(void* adj, uint32_t* one_launch) {
  Adj* a = (Adj*)adj;
  if (!a || !one_launch) return fail(ST_ERR_ARG, "st_adj_sor_path: null argument");
  *one_launch = a->lastOne ? 1u : 0u;
  return ST_OK;
}

}  // extern "C"

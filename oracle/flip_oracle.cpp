// flip_oracle.cpp - TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the reference's FLIP substep (the WGSL in src/flip/wgsl-core.js,
// wgsl-blocks.js, wgsl-solver.js, wgsl-narrowband.js and the SDF helpers of
// src/gpu.js:5-39), statement for statement, with each kernel of the reference a
// loop here.  It is the parity checker for the HIP library: only tests/, the smoke
// check and bench.py's CPU baseline may load it.  It is never the product path.
//
// Two instances: Real = float follows the reference in f32 (with contraction off,
// so every rounding is the one the source writes); Real = double runs the same
// algorithm in f64 to state the error of f32 results.
//
// Where the reference is order-dependent through atomics (free-stack pops, list
// order), this file picks one deterministic order: tiles in increasing tile
// number, the k-th new tile popping the k-th slot from the top of the stack.  The
// set of tiles with storage and every per-cell value do not depend on that order
// when the pool does not overflow.
//
// The ocean side boundary (oceanEta/oceanVel/layerW, wgsl-core.js:201-226) and its
// branches in classify, normBody, g2p, the narrow band and the whitewater are
// restated with the rest.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "splashtank.h"
#include "st_math.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const uint32_t NONE = 0xffffffffu;

template <class R> struct M;
template <> struct M<float> {
  static float exp(float x) { return st_expf(x); }
  static void sincos(float x, float* sn, float* cs) { st_sincosf(x, sn, cs); }
  static void sincosWide(float x, float* sn, float* cs) { st_sincosf_wide(x, sn, cs); }
  static float rnd(float x) { return rintf(x); }
  static float flo(float x) { return floorf(x); }
  static float sq(float x) { return sqrtf(x); }
  static float cei(float x) { return ceilf(x); }
};
template <> struct M<double> {
  static double exp(double x) { return std::exp(x); }
  static void sincos(double x, double* sn, double* cs) { *sn = std::sin(x); *cs = std::cos(x); }
  static void sincosWide(double x, double* sn, double* cs) { *sn = std::sin(x); *cs = std::cos(x); }
  static double rnd(double x) { return std::nearbyint(x); }
  static double flo(double x) { return std::floor(x); }
  static double sq(double x) { return std::sqrt(x); }
  static double cei(double x) { return std::ceil(x); }
};

template <class R> struct V3 {
  R x, y, z;
  R& operator[](int a) { return a == 0 ? x : (a == 1 ? y : z); }
  R operator[](int a) const { return a == 0 ? x : (a == 1 ? y : z); }
};
template <class R> V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> R len(V3<R> a) { return M<R>::sq(dot(a, a)); }
template <class R> R mixs(R a, R b, R t) { return a * (R(1) - t) + b * t; }
// WGSL leaves max/min of a NaN to the implementation; this restatement and the
// kernels both take IEEE maxNum/minNum (the non-NaN operand)
template <class R> R fmx(R a, R b) { return std::fmax(a, b); }
template <class R> R fmn(R a, R b) { return std::fmin(a, b); }
template <class R> R clampr(R x, R lo, R hi) { return fmn(fmx(x, lo), hi); }
template <class R> R smooth(R lo, R hi, R x) {
  R t = clampr((x - lo) / (hi - lo), R(0), R(1));
  return t * t * (R(3) - R(2) * t);
}
template <class R> int32_t toi(R x) {  // WGSL i32(): truncate, saturate
  if (!(x == x)) return 0;
  if (x >= R(2147483520.0)) return 2147483647;
  if (x <= R(-2147483648.0)) return INT32_MIN;
  return (int32_t)x;
}
template <class R> uint32_t tou(R x) {  // WGSL u32(): truncate, saturate
  if (!(x > R(0))) return 0u;
  if (x >= R(4294967040.0)) return 4294967295u;
  return (uint32_t)x;
}
inline int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }
inline int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }
inline int32_t iclamp(int32_t x, int32_t lo, int32_t hi) { return imin(imax(x, lo), hi); }

struct I3 { int32_t x, y, z; };

// the parameter block of wgsl-core.js:3-20, as packed by FlipSim.step (sim.js:196-231)
template <class R> struct Params {
  uint32_t n, nx, ny, nz, ncols, open, ncell;
  R dt, h, ih, gravity, flip, rest, comp, omega, x0, y0, z0, vmax, margin, scale, x1, y1, z1;
  R q0, q1, q2, q3, q4, q5, q6, q7, q8, q9, q10, q11, q12;
  uint32_t lcap, tx, ty, tz, scap;
  R sheetMin;
  uint32_t seed;
  uint32_t nb;
  R band;
  uint32_t ocean;
  uint32_t rk2, extrap;  // opt-in modes (0: the reference's path)
  uint32_t nbInit, nfill, nfv;
  R fills[16][4];
  // whitewater and air field (wgsl-core.js:10-16)
  R sprRate, sprMin, sprLife, mistFrac, sprDrag, mistDrag, vs;
  uint32_t smax;
  R wwFoamLife, wwRise, wwCrest, wwImpact;
  uint32_t airOn;
  R airCouple, airDecay, airH;
  uint32_t aNx, aNy, aNz;
  R airBlur;
  uint32_t ndrip;
  R dripRate, dripDry;
  // ocean (wgsl-core.js:12-13, 19; sim.js:207-218)
  uint32_t nwave;
  R otime, olevel, olayer, odepth, oy0;
  R waves[64][4];
};

template <class R> struct Sim {
  // sizes (sim.js:47-65)
  R d, h, ratio, rest;
  double hd_ = 0, dd_ = 0;  // h and d as the JS doubles they are in sim.js:48
  int32_t nx, ny, nz;
  uint32_t tx, ty, tz, ntiles, scap, scapMax, maxn, ncell, open;
  R bx0, by0, bz0, bx1, by1, bz1;
  uint32_t N = 0;
  std::vector<R> pos, vel;  // 4 per particle
  std::vector<uint32_t> tmap, need, slotTile, fstack, slotList, newList, lst;
  uint32_t ctl[16];
  std::vector<int32_t> cnt, accVi, accWi;
  std::vector<uint32_t> ctype;
  std::vector<R> pres, rhs, faceV, faceO, velG;
  uint32_t lcap = 0;
  bool velGOn = false;
  // colliders (gpu.js:42-70): 8 vec4 per entry, and the distance grids
  std::vector<R> cols, sdf;
  std::vector<float> cfPrev;
  uint64_t sdfSig = ~0ull, sdfSigPrev = ~0ull;
  bool haveCols = false, solidDirty = true;
  uint32_t ncolsTotal = 0;
  double tm = 0;  // seconds since start, a JS double (sim.js:200)
  uint32_t seed = 0;
  Params<R> P{};
  int iters = 1;
  bool nbOn = false;
  uint32_t initStep = 0;
  bool onlyTiles = false;
  // whitewater (setSprayCap, airAlloc; sim.js:333-351): 4 values per point / air cell
  uint32_t sprayCap = 0;
  std::vector<R> sPos, sVel, sAux, air, air2;
  std::vector<int32_t> airAcc;
  uint32_t sctl[4] = {0, 0, 0, 0};
  uint32_t airN = 0, aNx = 1, aNy = 1, aNz = 1;
  double airH = 1.0;
  bool airOnStep = false, sprOnStep = false;
  std::vector<R> drip;  // 8 per sample: local position, collider, local normal, wetness
  uint32_t ndrip = 0;
  bool dripOnStep = false;

  // Face values: from normFaces on, the reference keeps the new and old face
  // velocity as f32 bits in the same words that held the fixed-point sums
  // (wgsl-core.js:24-31).  The f32 instance does the same, so a face that no
  // normFaces pass converted (the spare slot) reads its integer bits as a float
  // exactly as the reference does; the f64 instance keeps separate arrays.
  static constexpr bool kF32 = std::is_same<R, float>::value;
  R fvR(uint32_t f) const { if constexpr (kF32) return st_u2f(uint32_t(accVi[f])); else return faceV[f]; }
  R foR(uint32_t f) const { if constexpr (kF32) return st_u2f(uint32_t(accWi[f])); else return faceO[f]; }
  void fvW(uint32_t f, R v) { if constexpr (kF32) accVi[f] = int32_t(st_f2u(float(v))); else faceV[f] = v; }
  void foW(uint32_t f, R v) { if constexpr (kF32) accWi[f] = int32_t(st_f2u(float(v))); else faceO[f] = v; }

  void makeSlots(uint32_t n) {
    const size_t C = 512;
    cnt.assign(n * C, 0);
    accVi.assign(n * C * 3, 0);
    accWi.assign(n * C * 3, 0);
    ctype.assign(n * C, 0u);
    pres.assign(n * C, R(0));
    rhs.assign(n * C, R(0));
    if (!kF32) {
      faceV.assign(n * C * 3, R(0));
      faceO.assign(n * C * 3, R(0));
    }
    velG.assign(n * C * 3, R(0));
    slotList.assign(n, 0u);
    slotTile.assign(n, 0u);
    fstack.assign(n, 0u);
    lcap = velGOn ? n * 512u : std::min<uint32_t>(n * 512u, maxn);
    lst.assign(size_t(lcap) * 2, 0u);
  }

  // grow (sim.js:137-152): a pool of n slots, the old slots' contents unchanged,
  // the new slots pushed on the free stack (growFree, wgsl-blocks.js:57-60; the
  // reference pushes them in atomic order, here in slot order, one of its orders)
  int grow(uint32_t n) {
    n = std::min(scapMax, std::max(n, scap + 1u));
    if (n <= scap) return ST_OK;
    const uint32_t old = scap;
    const size_t C = 512;
    cnt.resize(n * C, 0);
    accVi.resize(n * C * 3, 0);
    accWi.resize(n * C * 3, 0);
    ctype.resize(n * C, 0u);
    pres.resize(n * C, R(0));
    rhs.resize(n * C, R(0));
    if (!kF32) {
      faceV.resize(n * C * 3, R(0));
      faceO.resize(n * C * 3, R(0));
    }
    velG.resize(n * C * 3, R(0));
    slotList.resize(n, 0u);
    slotTile.resize(n, 0u);
    fstack.resize(n, 0u);
    for (uint32_t t = 0; t < n - old; t++) fstack[ctl[2] + t] = old + t;
    ctl[2] += n - old;
    scap = n;
    lcap = velGOn ? n * 512u : std::min<uint32_t>(n * 512u, maxn);
    lst.assign(size_t(lcap) * 2, 0u);  // the lists are rebuilt every step
    return ST_OK;
  }

  int init(const st_opts* o) {
    d = R(o->d);
    double rt = o->ratio > 0 ? o->ratio : 2.0;
    rt = std::max(1.4, std::min(4.0, rt));
    ratio = R(rt);
    h = R(rt * o->d);  // sim.js:48: h = ratio * d in JS doubles
    double hd = rt * o->d;
    hd_ = hd; dd_ = o->d;
    rest = R(rt * rt * rt);
    double sz[3] = {o->size[0] > 0 ? o->size[0] : 3.0, o->size[1] > 0 ? o->size[1] : 2.6, o->size[2] > 0 ? o->size[2] : 3.0};
    nx = std::max(4, (int)std::floor(sz[0] / hd + 0.5));  // Math.round
    ny = std::max(4, (int)std::floor(sz[1] / hd + 0.5));
    nz = std::max(4, (int)std::floor(sz[2] / hd + 0.5));
    double W = nx * hd, H = ny * hd, D = nz * hd;
    bx0 = R(-W / 2); bx1 = R(W / 2); by0 = R(0); by1 = R(H); bz0 = R(-D / 2); bz1 = R(D / 2);
    open = o->open;
    maxn = o->max_n;
    ncell = uint32_t(nx) * ny * nz;
    tx = (nx + 1 + 7) / 8; ty = (ny + 1 + 7) / 8; tz = (nz + 1 + 7) / 8;
    ntiles = tx * ty * tz;
    scapMax = ntiles + 1;
    scap = o->slots ? std::max<uint32_t>(2, std::min(scapMax, o->slots)) : scapMax;
    pos.assign(size_t(maxn) * 4, R(0));
    vel.assign(size_t(maxn) * 4, R(0));
    tmap.assign(ntiles, NONE);
    need.assign(ntiles, 0u);
    newList.assign(size_t(ntiles) * 2, 0u);
    makeSlots(scap);
    for (uint32_t k = 0; k + 1 < scap; k++) fstack[k] = scap - 1 - k;  // sim.js:96
    std::memset(ctl, 0, sizeof ctl);
    ctl[2] = scap - 1;
    return ST_OK;
  }

  // ---- helpers of wgsl-core.js ----
  uint32_t tileOf(int32_t i, int32_t j, int32_t k) const {  // wgsl-core.js:78
    return uint32_t(i >> 3) + P.tx * (uint32_t(j >> 3) + P.ty * uint32_t(k >> 3));
  }
  uint32_t cellIdx(int32_t i, int32_t j, int32_t k) const {  // wgsl-core.js:80-83
    uint32_t s = tmap[tileOf(i, j, k)];
    return (s == NONE ? 0u : s) * 512u + uint32_t(i & 7) + 8u * (uint32_t(j & 7) + 8u * uint32_t(k & 7));
  }
  uint32_t faceIndex(uint32_t axis, int32_t i, int32_t j, int32_t k) const {  // wgsl-core.js:84-89
    int32_t X = int32_t(P.nx), Y = int32_t(P.ny), Z = int32_t(P.nz);
    if (axis == 0) return cellIdx(iclamp(i, 0, X), iclamp(j, 0, Y - 1), iclamp(k, 0, Z - 1)) * 3u;
    if (axis == 1) return cellIdx(iclamp(i, 0, X - 1), iclamp(j, 0, Y), iclamp(k, 0, Z - 1)) * 3u + 1u;
    return cellIdx(iclamp(i, 0, X - 1), iclamp(j, 0, Y - 1), iclamp(k, 0, Z)) * 3u + 2u;
  }
  I3 slotCoord(uint32_t s, uint32_t local) const {  // wgsl-core.js:90-93
    uint32_t t = slotTile[s], tt = t / P.tx;
    return I3{int32_t((t % P.tx) * 8u + (local & 7u)), int32_t((tt % P.ty) * 8u + ((local >> 3) & 7u)),
              int32_t((tt / P.ty) * 8u + (local >> 6))};
  }
  uint32_t ctypeOf(int32_t i, int32_t j, int32_t k) const {  // wgsl-core.js:104-112
    if (i < 0) return (P.open & 1u) ? 0u : 2u;
    if (i >= int32_t(P.nx)) return (P.open & 2u) ? 0u : 2u;
    if (j < 0) return (P.open & 4u) ? 0u : 2u;
    if (j >= int32_t(P.ny)) return (P.open & 8u) ? 0u : 2u;
    if (k < 0) return (P.open & 16u) ? 0u : 2u;
    if (k >= int32_t(P.nz)) return (P.open & 32u) ? 0u : 2u;
    return ctype[cellIdx(i, j, k)] & 0xffu;
  }
  V3<R> colRow(uint32_t k, uint32_t r) const { const R* c = &cols[(k * 8 + r) * 4]; return v3(c[0], c[1], c[2]); }
  R colW(uint32_t k, uint32_t r) const { return cols[(k * 8 + r) * 4 + 3]; }
  // sdfEval (gpu.js:6-33): trilinear distance and world normal
  void sdfEval(uint32_t k, V3<R> p, R& phiOut, V3<R>& wnOut) const {
    const R* a = &cols[(k * 8) * 4];
    const R* b = &cols[(k * 8 + 1) * 4];
    V3<R> r0 = colRow(k, 2), r1 = colRow(k, 3), r2 = colRow(k, 4);
    V3<R> dd = p - colRow(k, 5);
    V3<R> lp = v3(dot(r0, dd), dot(r1, dd), dot(r2, dd));
    V3<R> g = (lp - v3(b[0], b[1], b[2])) / b[3];
    V3<R> dims = v3(a[1], a[2], a[3]);
    R e = R(1.0001);
    V3<R> gc = v3(clampr(g.x, R(0), dims.x - e), clampr(g.y, R(0), dims.y - e), clampr(g.z, R(0), dims.z - e));
    V3<R> fl = v3(M<R>::flo(gc.x), M<R>::flo(gc.y), M<R>::flo(gc.z));
    V3<R> f = gc - fl;
    uint32_t snx = tou(a[1]), sny = tou(a[2]);
    uint32_t base = tou(a[0]) + tou(fl.x) + snx * (tou(fl.y) + sny * tou(fl.z));
    uint32_t s = snx * sny;
    R v000 = sdf[base], v100 = sdf[base + 1], v010 = sdf[base + snx], v110 = sdf[base + snx + 1];
    R v001 = sdf[base + s], v101 = sdf[base + s + 1], v011 = sdf[base + s + snx], v111 = sdf[base + s + snx + 1];
    R c00 = mixs(v000, v100, f.x), c10 = mixs(v010, v110, f.x), c01 = mixs(v001, v101, f.x), c11 = mixs(v011, v111, f.x);
    R c0 = mixs(c00, c10, f.y), c1 = mixs(c01, c11, f.y);
    R phi = mixs(c0, c1, f.z);
    V3<R> out = g - gc;
    V3<R> gl;
    if (dot(out, out) > R(0)) {
      R l = len(out);
      phi = phi + l * b[3];
      gl = out / l;
    } else {
      gl = v3(mixs(mixs(v100 - v000, v110 - v010, f.y), mixs(v101 - v001, v111 - v011, f.y), f.z),
              mixs(c10 - c00, c11 - c01, f.z), c1 - c0);
      R l = len(gl);
      gl = gl / fmx<R>(l, R(1e-9));
    }
    wnOut = r0 * gl.x + r1 * gl.y + r2 * gl.z;
    phiOut = phi;
  }
  V3<R> colSurfVel(uint32_t k, V3<R> p) const {  // gpu.js:35-38
    V3<R> c = colRow(k, 5), w = colRow(k, 7), v = colRow(k, 6);
    V3<R> q = p - c;
    V3<R> cr = v3(w.y * q.z - w.z * q.y, w.z * q.x - w.x * q.z, w.x * q.y - w.y * q.x);
    return v + cr;
  }
  bool colNear(uint32_t k, V3<R> p) const {  // wgsl-core.js:114-122
    if (P.q10 > R(0.5)) return true;
    const R* b = &cols[(k * 8 + 1) * 4];
    V3<R> dd = p - colRow(k, 5);
    V3<R> lp = v3(dot(colRow(k, 2), dd), dot(colRow(k, 3), dd), dot(colRow(k, 4), dd));
    V3<R> g = (lp - v3(b[0], b[1], b[2])) / b[3];
    V3<R> dims = v3(cols[k * 32 + 1], cols[k * 32 + 2], cols[k * 32 + 3]);
    R lo = R(-1.0) - P.q0 / b[3];
    R hq = P.q0 / b[3];
    return g.x > lo && g.y > lo && g.z > lo && g.x < dims.x + hq && g.y < dims.y + hq && g.z < dims.z + hq;
  }
  int32_t colAt(V3<R> p) const {  // wgsl-core.js:123-126
    for (uint32_t k = 0; k < P.ncols; k++) {
      if (!colNear(k, p)) continue;
      R phi; V3<R> wn;
      sdfEval(k, p, phi, wn);
      if (phi < P.q0) return int32_t(k);
    }
    return -1;
  }
  V3<R> cellCenter(int32_t i, int32_t j, int32_t k) const {  // wgsl-core.js:127
    return v3(P.x0 + (R(i) + R(0.5)) * P.h, P.y0 + (R(j) + R(0.5)) * P.h, P.z0 + (R(k) + R(0.5)) * P.h);
  }
  int32_t colAtCell(int32_t i, int32_t j, int32_t k) const {  // wgsl-core.js:128-131
    if (i < 0 || j < 0 || k < 0 || i >= int32_t(P.nx) || j >= int32_t(P.ny) || k >= int32_t(P.nz)) return -1;
    return int32_t((ctype[cellIdx(i, j, k)] >> 8) & 0xffu) - 1;
  }
  I3 particleCell(V3<R> p) const {  // wgsl-core.js:132-134
    I3 c{toi(M<R>::flo((p.x - P.x0) * P.ih)), toi(M<R>::flo((p.y - P.y0) * P.ih)), toi(M<R>::flo((p.z - P.z0) * P.ih))};
    return I3{iclamp(c.x, 0, int32_t(P.nx) - 1), iclamp(c.y, 0, int32_t(P.ny) - 1), iclamp(c.z, 0, int32_t(P.nz) - 1)};
  }
  bool hasStorage(I3 c) const { return tmap[tileOf(c.x, c.y, c.z)] != NONE; }
  // sampleBoth (wgsl-core.js:138-154)
  void sampleBoth(V3<R> p, V3<R>& vn, V3<R>& vo) const {
    for (uint32_t axis = 0; axis < 3; axis++) {
      V3<R> g = v3((p.x - P.x0) * P.ih, (p.y - P.y0) * P.ih, (p.z - P.z0) * P.ih);
      if (axis == 0) { g.y = g.y - R(0.5); g.z = g.z - R(0.5); }
      else if (axis == 1) { g.x = g.x - R(0.5); g.z = g.z - R(0.5); }
      else { g.x = g.x - R(0.5); g.y = g.y - R(0.5); }
      V3<R> f0 = v3(M<R>::flo(g.x), M<R>::flo(g.y), M<R>::flo(g.z));
      V3<R> fr = g - f0;
      I3 i0{toi(f0.x), toi(f0.y), toi(f0.z)};
      R sn = R(0), so = R(0);
      for (uint32_t c = 0; c < 8; c++) {
        int32_t dx = int32_t(c & 1u), dy = int32_t((c >> 1) & 1u), dz = int32_t((c >> 2) & 1u);
        R w = (dx == 1 ? fr.x : R(1) - fr.x) * (dy == 1 ? fr.y : R(1) - fr.y) * (dz == 1 ? fr.z : R(1) - fr.z);
        uint32_t f = faceIndex(axis, i0.x + dx, i0.y + dy, i0.z + dz);
        sn = sn + w * fvR(f);
        so = so + w * foR(f);
      }
      vn[int(axis)] = sn;
      vo[int(axis)] = so;
    }
  }
  // narrow band helpers (wgsl-core.js:160-200)
  uint32_t colBits(uint32_t c) const { return (ctype[c] >> 8) & 0xffu; }
  R phiOf(uint32_t c) const { return R(8.0) - R(ctype[c] >> 16) / R(4096.0); }
  void phiPut(uint32_t c, R v) {
    uint32_t q = tou(clampr(M<R>::rnd((R(8.0) - v) * R(4096.0)), R(0.0), R(65535.0)));
    ctype[c] = (ctype[c] & 0xffffu) | (q << 16);
  }
  bool inDom(I3 p) const { return p.x >= 0 && p.y >= 0 && p.z >= 0 && p.x < int32_t(P.nx) && p.y < int32_t(P.ny) && p.z < int32_t(P.nz); }
  V3<R> facePos(I3 p, uint32_t axis) const {
    return v3(P.x0 + (R(p.x) + (axis == 0 ? R(0) : R(0.5))) * P.h, P.y0 + (R(p.y) + (axis == 1 ? R(0) : R(0.5))) * P.h,
              P.z0 + (R(p.z) + (axis == 2 ? R(0) : R(0.5))) * P.h);
  }
  R phiSample(V3<R> x) const {  // wgsl-core.js:169-183
    V3<R> g = v3((x.x - P.x0) * P.ih - R(0.5), (x.y - P.y0) * P.ih - R(0.5), (x.z - P.z0) * P.ih - R(0.5));
    V3<R> gc = v3(clampr(g.x, R(0), R(P.nx) - R(1.001)), clampr(g.y, R(0), R(P.ny) - R(1.001)), clampr(g.z, R(0), R(P.nz) - R(1.001)));
    V3<R> f0 = v3(M<R>::flo(gc.x), M<R>::flo(gc.y), M<R>::flo(gc.z));
    V3<R> fr = gc - f0;
    I3 i0{toi(f0.x), toi(f0.y), toi(f0.z)};
    R s = R(0), ws = R(0);
    for (uint32_t c = 0; c < 8; c++) {
      int32_t dx = int32_t(c & 1u), dy = int32_t((c >> 1) & 1u), dz = int32_t((c >> 2) & 1u);
      R w = (dx == 1 ? fr.x : R(1) - fr.x) * (dy == 1 ? fr.y : R(1) - fr.y) * (dz == 1 ? fr.z : R(1) - fr.z);
      uint32_t ci = cellIdx(i0.x + dx, i0.y + dy, i0.z + dz);
      if (colBits(ci) != 0u || w <= R(0)) continue;
      s = s + w * phiOf(ci);
      ws = ws + w;
    }
    if (ws < R(1.0e-4)) return R(0.5);
    return s / ws;
  }
  R velGAxis(V3<R> x, uint32_t axis) const {  // wgsl-core.js:185-196
    V3<R> g = v3((x.x - P.x0) * P.ih, (x.y - P.y0) * P.ih, (x.z - P.z0) * P.ih);
    if (axis == 0) { g.y = g.y - R(0.5); g.z = g.z - R(0.5); }
    else if (axis == 1) { g.x = g.x - R(0.5); g.z = g.z - R(0.5); }
    else { g.x = g.x - R(0.5); g.y = g.y - R(0.5); }
    V3<R> f0 = v3(M<R>::flo(g.x), M<R>::flo(g.y), M<R>::flo(g.z));
    V3<R> fr = g - f0;
    I3 i0{toi(f0.x), toi(f0.y), toi(f0.z)};
    R s = R(0);
    for (uint32_t c = 0; c < 8; c++) {
      int32_t dx = int32_t(c & 1u), dy = int32_t((c >> 1) & 1u), dz = int32_t((c >> 2) & 1u);
      R w = (dx == 1 ? fr.x : R(1) - fr.x) * (dy == 1 ? fr.y : R(1) - fr.y) * (dz == 1 ? fr.z : R(1) - fr.z);
      s = s + w * velG[faceIndex(axis, i0.x + dx, i0.y + dy, i0.z + dz)];
    }
    return s;
  }
  V3<R> velGVec(V3<R> x) const { return v3(velGAxis(x, 0), velGAxis(x, 1), velGAxis(x, 2)); }
  bool calmV(V3<R> v) const { return len(v) < R(9.0) * P.h; }
  R oceanEta(R x, R z) const {  // wgsl-core.js:202-206
    R e = R(0);
    for (uint32_t k = 0; k < P.nwave; k++) {
      const R* A = P.waves[k * 2];
      const R* B = P.waves[k * 2 + 1];
      R sn, cs;
      M<R>::sincosWide(A[0] * x + A[1] * z - A[3] * P.otime + B[0], &sn, &cs);
      e = e + A[2] * cs;
    }
    return P.olevel + e;
  }
  V3<R> oceanVel(V3<R> p) const {  // wgsl-core.js:208-220
    const R D = fmx<R>(P.odepth, R(1.0e-3));
    const R eta = oceanEta(p.x, p.z);
    const R ys = clampr((p.y - P.oy0) * D / fmx<R>(D + eta - P.olevel, R(0.1) * D), R(0), D);
    V3<R> v = v3(R(0), R(0), R(0));
    for (uint32_t k = 0; k < P.nwave; k++) {
      const R* A = P.waves[k * 2];
      const R* B = P.waves[k * 2 + 1];
      const R kk = fmx<R>(B[1], R(1.0e-4));
      R sn, cs;
      M<R>::sincosWide(A[0] * p.x + A[1] * p.z - A[3] * P.otime + B[0], &sn, &cs);
      const R e1 = M<R>::exp(kk * (ys - D));
      const R e2 = M<R>::exp(-kk * (ys + D));
      const R dn = fmx<R>(R(1) - M<R>::exp(R(-2.0) * kk * D), R(1.0e-4));
      const R aw = A[2] * A[3];
      const R hc = aw * (e1 + e2) / dn * cs / kk;
      v = v + v3(hc * A[0], aw * (e1 - e2) / dn * sn, hc * A[1]);
    }
    return v;
  }
  R layerW(V3<R> p) const {  // wgsl-core.js:222-226
    if (P.ocean == 0u) return R(0);
    const R dw = fmn<R>(fmn<R>(p.x - P.x0, P.x1 - p.x), fmn<R>(p.z - P.z0, P.z1 - p.z));
    return clampr(R(1) - dw / fmx<R>(P.olayer, R(1.0e-4)), R(0), R(1));
  }
  R fillPhi(V3<R> x) const {  // wgsl-core.js:228-238
    R dd = R(1.0e9);
    for (uint32_t k = 0; k < P.nfill; k++) {
      const R* A = P.fills[k * 2];
      const R* B = P.fills[k * 2 + 1];
      if (A[0] < R(1.5)) {
        V3<R> q = v3(fmx<R>(A[1] - x.x, x.x - B[0]), fmx<R>(A[2] - x.y, x.y - B[1]), fmx<R>(A[3] - x.z, x.z - B[2]));
        V3<R> qp = v3(fmx<R>(q.x, R(0)), fmx<R>(q.y, R(0)), fmx<R>(q.z, R(0)));
        dd = fmn<R>(dd, len(qp) + fmn<R>(fmx<R>(q.x, fmx<R>(q.y, q.z)), R(0)));
      } else {
        dd = fmn<R>(dd, len(x - v3(A[1], A[2], A[3])) - B[0]);
      }
    }
    for (uint32_t k = P.ncols; k < P.ncols + P.nfv; k++) {
      R phi; V3<R> wn;
      sdfEval(k, x, phi, wn);
      dd = fmn<R>(dd, phi);
    }
    if (P.ocean != 0u) dd = fmn<R>(dd, x.y - oceanEta(x.x, x.z));
    return dd * P.ih;
  }
  void markAround(I3 cc) {  // wgsl-core.js:239-246
    I3 lo{imax(cc.x - 2, 0) >> 3, imax(cc.y - 2, 0) >> 3, imax(cc.z - 2, 0) >> 3};
    I3 hi{imin(cc.x + 2, int32_t(P.tx) * 8 - 1) >> 3, imin(cc.y + 2, int32_t(P.ty) * 8 - 1) >> 3, imin(cc.z + 2, int32_t(P.tz) * 8 - 1) >> 3};
    for (int32_t tk = lo.z; tk <= hi.z; tk++)
      for (int32_t tj = lo.y; tj <= hi.y; tj++)
        for (int32_t ti = lo.x; ti <= hi.x; ti++)
          __atomic_store_n(&need[uint32_t(ti) + P.tx * (uint32_t(tj) + P.ty * uint32_t(tk))], 1u, __ATOMIC_RELAXED);
  }
  template <class F> void forActive(F f) {  // acell (wgsl-core.js:96-103), one call per active cell
    uint32_t sc = ctl[4];
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < int64_t(sc); b++) {
      uint32_t s = slotList[b];
      for (uint32_t local = 0; local < 512u; local++) f(s * 512u + local, slotCoord(s, local));
    }
  }

  // ---- blocks (wgsl-blocks.js:4-60) ----
  void stageBlocks() {
    std::fill(need.begin(), need.end(), 0u);  // tileNeedClear
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < int64_t(P.n); ii++) {  // tileMark
      size_t i = size_t(ii);
      if (pos[i * 4 + 1] < R(-1.0e5)) continue;
      markAround(particleCell(v3(pos[i * 4], pos[i * 4 + 1], pos[i * 4 + 2])));
    }
    if (nbOn && initStep != 1) {  // tileMarkPhi (wgsl-narrowband.js:6-10), last step's blocks
      uint32_t sc = ctl[4];
      for (uint32_t b = 0; b < sc; b++) {
        uint32_t s = slotList[b];
        for (uint32_t local = 0; local < 512u; local++) {
          I3 p = slotCoord(s, local);
          if (!inDom(p) || phiOf(s * 512u + local) >= R(0.5)) continue;
          markAround(p);
        }
      }
    }
    if (initStep) {  // tileMarkFill (wgsl-narrowband.js:13-18)
      for (uint32_t t = 0; t < ntiles; t++) {
        uint32_t tt = t / P.tx;
        V3<R> tc = v3(R(t % P.tx), R(tt % P.ty), R(tt / P.ty)) * R(8.0);
        tc = tc + v3(R(4.0), R(4.0), R(4.0));
        V3<R> q = v3(P.x0, P.y0, P.z0) + tc * P.h;
        if (fillPhi(q) < R(9.0)) need[t] = 1u;
      }
    }
    ctl[3] = 0; ctl[4] = 0; ctl[5] = 0; ctl[6] = 0;  // tileReset
    for (uint32_t t = 0; t < ntiles; t++) {        // tileFree, in tile order
      uint32_t nd = need[t], s = tmap[t];
      if (nd == 0u && s != NONE) { tmap[t] = NONE; fstack[ctl[2]++] = s; }
      else if (nd != 0u && s == NONE) { uint32_t n = ctl[3]++; newList[n * 2] = t; newList[n * 2 + 1] = NONE; }
    }
    uint32_t top = ctl[2];                            // tileAlloc: new tile k pops stack entry top-1-k
    for (uint32_t k = 0; k < ctl[3]; k++) {
      uint32_t old = top - k;
      if (old >= 1u && old <= P.scap) {
        uint32_t s = fstack[old - 1], t = newList[k * 2];
        tmap[t] = s; slotTile[s] = t; newList[k * 2 + 1] = s;
      } else {
        ctl[5]++;
      }
    }
    ctl[2] = top - ctl[3];
    if (ctl[2] > P.scap) ctl[2] = 0;  // tileFix
    uint32_t nn = ctl[3];              // initNew
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < int64_t(nn); k++) {
      uint32_t s = newList[k * 2 + 1];
      if (s == NONE) continue;
      uint32_t tl = newList[k * 2], tt = tl / P.tx;
      for (uint32_t local = 0; local < 512u; local++) {
        I3 p{int32_t((tl % P.tx) * 8u + (local & 7u)), int32_t((tt % P.ty) * 8u + ((local >> 3) & 7u)), int32_t((tt / P.ty) * 8u + (local >> 6))};
        uint32_t c = s * 512u + local;
        pres[c] = R(0);
        ctype[c] = uint32_t(colAt(cellCenter(p.x, p.y, p.z)) + 1) << 8;
        velG[c * 3] = R(0); velG[c * 3 + 1] = R(0); velG[c * 3 + 2] = R(0);
      }
    }
    for (uint32_t t = 0; t < ntiles; t++) {  // slotListBuild, in tile order
      uint32_t s = tmap[t];
      if (s != NONE) slotList[ctl[4]++] = s;
    }
    if (initStep) phiInitAll();
  }
  void phiInitAll() {  // wgsl-narrowband.js:21-35
    forActive([&](uint32_t c, I3 p) {
      R ph = clampr(fillPhi(cellCenter(p.x, p.y, p.z)), R(-7.9), R(7.9));
      if (ph < R(0)) ph = fmn<R>(ph, R(-1.5));
      if (!inDom(p)) ph = R(8.0);
      if (P.nbInit == 2u) ph = fmn<R>(ph, phiOf(c));
      phiPut(c, ph);
      if (P.nbInit == 1u)
        for (uint32_t ax = 0; ax < 3; ax++) {
          V3<R> fp = facePos(p, ax);
          R v = R(0);
          if (P.ocean != 0u && fp.y < oceanEta(fp.x, fp.z)) v = oceanVel(fp)[int(ax)];
          velG[c * 3 + ax] = v;
        }
    });
  }
  // ---- solver (wgsl-solver.js) ----
  void stageClear() {  // clearSlots 5-15
    uint32_t sc = ctl[4];
    for (uint32_t blk = 0; blk <= sc; blk++) {
      uint32_t s = blk < sc ? slotList[blk] : 0u;
      for (uint32_t l = 0; l < 512u; l++) {
        uint32_t c = s * 512u + l;
        cnt[c] = 0;
        for (int a = 0; a < 3; a++) { accVi[c * 3 + a] = 0; accWi[c * 3 + a] = 0; }
        if (s == 0u) rhs[c] = R(8.0);
      }
    }
  }
  void stageP2G() {  // p2g 16-40
    uint32_t vmaxBits = 0;
    int32_t* AV = accVi.data();
    int32_t* AW = accWi.data();
    int32_t* CN = cnt.data();
#pragma omp parallel for schedule(static) reduction(max : vmaxBits)
    for (int64_t ii = 0; ii < int64_t(P.n); ii++) {
      size_t i = size_t(ii);
      V3<R> p = v3(pos[i * 4], pos[i * 4 + 1], pos[i * 4 + 2]);
      if (pos[i * 4 + 1] < R(-1.0e5)) continue;
      I3 cc = particleCell(p);
      if (!hasStorage(cc)) continue;
      V3<R> v = v3(vel[i * 4], vel[i * 4 + 1], vel[i * 4 + 2]);
      float lv = float(len(v));  // bitcast<u32>(length(v)); f32 bits of the f32 length
      uint32_t bits = st_f2u(lv);
      if (bits > vmaxBits) vmaxBits = bits;
      for (uint32_t axis = 0; axis < 3; axis++) {
        V3<R> q = v3((p.x - P.x0) * P.ih, (p.y - P.y0) * P.ih, (p.z - P.z0) * P.ih);
        if (axis == 0) { q.y = q.y - R(0.5); q.z = q.z - R(0.5); }
        else if (axis == 1) { q.x = q.x - R(0.5); q.z = q.z - R(0.5); }
        else { q.x = q.x - R(0.5); q.y = q.y - R(0.5); }
        V3<R> f0 = v3(M<R>::flo(q.x), M<R>::flo(q.y), M<R>::flo(q.z));
        V3<R> fr = q - f0;
        I3 i0{toi(f0.x), toi(f0.y), toi(f0.z)};
        R vc = v[int(axis)];
        for (uint32_t c = 0; c < 8; c++) {
          int32_t dx = int32_t(c & 1u), dy = int32_t((c >> 1) & 1u), dz = int32_t((c >> 2) & 1u);
          R w = (dx == 1 ? fr.x : R(1) - fr.x) * (dy == 1 ? fr.y : R(1) - fr.y) * (dz == 1 ? fr.z : R(1) - fr.z);
          if (w <= R(0)) continue;
          uint32_t idx = faceIndex(axis, i0.x + dx, i0.y + dy, i0.z + dz);
          int32_t av = toi(M<R>::rnd(w * vc * P.scale));
          int32_t aw = toi(M<R>::rnd(w * P.scale));
#pragma omp atomic
          AV[idx] += av;
#pragma omp atomic
          AW[idx] += aw;
        }
      }
      uint32_t ci = cellIdx(cc.x, cc.y, cc.z);
#pragma omp atomic
      CN[ci] += 1;
    }
    ctl[6] = std::max(ctl[6], vmaxBits);
  }
  void solidBuild() {  // 41-45
    forActive([&](uint32_t c, I3 p) {
      ctype[c] = (ctype[c] & 0xffff00ffu) | (uint32_t(colAt(cellCenter(p.x, p.y, p.z)) + 1) << 8);
    });
  }
  void classify() {  // 46-74
    forActive([&](uint32_t c, I3 p) {
      uint32_t ty = 0;
      uint32_t sb = ctype[c] & 0xffffff00u;
      int32_t n = cnt[c];
      V3<R> cx = cellCenter(p.x, p.y, p.z);
      if (!inDom(p)) ty = 0;
      else if ((sb & 0xff00u) != 0u) ty = 2;
      else if (P.ocean != 0u && layerW(cx) >= R(0.5)) ty = cx.y < oceanEta(cx.x, cx.z) ? 2u : 0u;
      else if (P.nb != 0u && phiOf(c) < R(-1.0) && (phiOf(c) < -P.band || calmV(velGVec(cx)))) ty = 1;
      else if (n > 0) {
        ty = 1;
        if (R(n) < P.sheetMin) {
          int32_t dth = toi(M<R>::cei(R(0.6) * P.rest));
          bool dense = false;
          for (uint32_t q = 0; q < 6; q++) {
            I3 r = p;
            if (q == 0) r.x -= 1; else if (q == 1) r.x += 1; else if (q == 2) r.y -= 1; else if (q == 3) r.y += 1; else if (q == 4) r.z -= 1; else r.z += 1;
            if (r.x < 0 || r.y < 0 || r.z < 0 || r.x >= int32_t(P.nx) || r.y >= int32_t(P.ny) || r.z >= int32_t(P.nz)) continue;
            if (cnt[cellIdx(r.x, r.y, r.z)] >= dth) dense = true;
          }
          if (!dense) ty = 0;
        }
      }
      ctype[c] = ty | sb;
      if (ty != 1u) pres[c] = R(0);
    });
  }
  void buildList() {  // listClear, buildList, listArgs (149-160), in slotList order
    ctl[0] = 0; ctl[1] = 0;
    uint32_t sc = ctl[4];
    for (uint32_t b = 0; b < sc; b++) {
      uint32_t s = slotList[b];
      for (uint32_t local = 0; local < 512u; local++) {
        uint32_t c = s * 512u + local;
        if ((ctype[c] & 0xffu) != 1u) continue;
        I3 p = slotCoord(s, local);
        uint32_t par = uint32_t((p.x + p.y + p.z) & 1);
        uint32_t k = ctl[par]++;
        if (k < lcap) lst[par * lcap + k] = c;
      }
    }
  }
  void nbVel() {  // wgsl-narrowband.js:71-85
    forActive([&](uint32_t c, I3 p) {
      if (!inDom(p)) return;
      if ((ctype[c] & 0xffu) != 1u) return;
      for (uint32_t ax = 0; ax < 3; ax++) {
        I3 lo = p;
        if (ax == 0) lo.x -= 1; else if (ax == 1) lo.y -= 1; else lo.z -= 1;
        if (!inDom(lo)) continue;
        if ((ctype[cellIdx(lo.x, lo.y, lo.z)] & 0xffu) != 1u) continue;
        uint32_t f = c * 3 + ax;
        if (accWi[f] > 0) continue;
        V3<R> fp = facePos(p, ax);
        R vc = velGAxis(fp - velGVec(fp) * P.dt, ax);
        accVi[f] = toi(M<R>::rnd(vc * P.scale));
        accWi[f] = toi(P.scale);
      }
    });
  }
  void stageClassify() {
    if (solidDirty) { solidBuild(); solidDirty = false; }
    classify();
    buildList();
    if (nbOn) nbVel();
  }
  void normFaces() {  // normBody / normFaces 75-105
    forActive([&](uint32_t c, I3 at) {
      for (uint32_t axis = 0; axis < 3; axis++) {
        uint32_t f = c * 3 + axis;
        I3 lo = at, hi = at;
        if (axis == 0) lo.x -= 1; else if (axis == 1) lo.y -= 1; else lo.z -= 1;
        uint32_t tl = ctypeOf(lo.x, lo.y, lo.z), th = ctypeOf(hi.x, hi.y, hi.z);
        int32_t ws = accWi[f], vs = accVi[f];
        R v = ws > 0 ? R(vs) / R(imax(ws, 1)) : R(0);
        R old = v;
        if (axis == 1) v = v - P.gravity * P.dt;
        if (tl == 2u || th == 2u) {
          R sv = R(0);
          int32_t ci = -1;
          if (tl == 2u) ci = colAtCell(lo.x, lo.y, lo.z);
          if (ci < 0 && th == 2u) ci = colAtCell(hi.x, hi.y, hi.z);
          V3<R> fp = facePos(at, axis);
          if (ci >= 0) sv = colSurfVel(uint32_t(ci), fp)[int(axis)];
          else if (P.ocean != 0u && ((tl == 2u && inDom(lo)) || (th == 2u && inDom(hi)))) sv = oceanVel(fp)[int(axis)];
          v = sv; old = sv;
        }
        foW(f, old);
        fvW(f, v);
      }
    });
  }
  void divergence() {  // 106-123
    forActive([&](uint32_t c, I3 a) {
      if ((ctype[c] & 0xffu) != 1u) { rhs[c] = R(0); return; }
      int32_t i = a.x, j = a.y, k = a.z;
      R div = (fvR(faceIndex(0, i + 1, j, k)) - fvR(faceIndex(0, i, j, k)) + fvR(faceIndex(1, i, j + 1, k)) -
               fvR(faceIndex(1, i, j, k)) + fvR(faceIndex(2, i, j, k + 1)) - fvR(faceIndex(2, i, j, k))) * P.ih;
      bool interior = true;
      if (ctypeOf(i - 1, j, k) == 0u || ctypeOf(i + 1, j, k) == 0u || ctypeOf(i, j - 1, k) == 0u || ctypeOf(i, j + 1, k) == 0u ||
          ctypeOf(i, j, k - 1) == 0u || ctypeOf(i, j, k + 1) == 0u)
        interior = false;
      R dens = R(cnt[c]) / P.rest - R(1.0);
      R tgt = P.comp * (interior ? dens : fmx<R>(dens, R(0)));
      if (P.nb != 0u && phiOf(c) < R(-1.0)) tgt = R(0);
      rhs[c] = (tgt - div) * P.h * P.h / P.dt;
    });
  }
  void relax(uint32_t c, uint32_t color) {  // 124-140
    I3 p = slotCoord(c >> 9, c & 511u);
    int32_t i = p.x, j = p.y, k = p.z;
    if (uint32_t((i + j + k) & 1) != color) return;
    if ((ctype[c] & 0xffu) != 1u) return;
    R sum = R(0), nns = R(0);
    for (uint32_t a = 0; a < 6; a++) {
      I3 q{i, j, k};
      if (a == 0) q.x -= 1; else if (a == 1) q.x += 1; else if (a == 2) q.y -= 1; else if (a == 3) q.y += 1; else if (a == 4) q.z -= 1; else q.z += 1;
      uint32_t ty = ctypeOf(q.x, q.y, q.z);
      if (ty == 2u) continue;
      nns = nns + R(1.0);
      if (ty == 1u) sum = sum + pres[cellIdx(q.x, q.y, q.z)];
    }
    if (nns < R(0.5)) { pres[c] = R(0); return; }
    R pg = (rhs[c] + sum) / nns;
    pres[c] = pres[c] + P.omega * (pg - pres[c]);
  }
  void stagePressure() {  // rbRed / rbBlack 161-164, loop at sim.js:291-293
    for (int it = 0; it < iters; it++) {
      for (uint32_t color = 0; color < 2; color++) {
        uint32_t n = std::min(ctl[color], lcap);
        const uint32_t* L = &lst[color * lcap];
#pragma omp parallel for schedule(static)
        for (int64_t t = 0; t < int64_t(n); t++) relax(L[t], color);
      }
    }
  }
  void project() {  // projBody / project 165-180
    forActive([&](uint32_t c, I3 at) {
      for (uint32_t axis = 0; axis < 3; axis++) {
        uint32_t f = c * 3 + axis;
        I3 lo = at, hi = at;
        if (axis == 0) lo.x -= 1; else if (axis == 1) lo.y -= 1; else lo.z -= 1;
        uint32_t tl = ctypeOf(lo.x, lo.y, lo.z), th = ctypeOf(hi.x, hi.y, hi.z);
        if (tl == 2u || th == 2u) continue;
        if (tl != 1u && th != 1u) continue;
        R pl = R(0), ph = R(0);
        if (tl == 1u) pl = pres[cellIdx(lo.x, lo.y, lo.z)];
        if (th == 1u) ph = pres[cellIdx(hi.x, hi.y, hi.z)];
        fvW(f, fvR(f) - P.dt * P.ih * (ph - pl));
      }
    });
  }
  // Opt-in velocity extrapolation after project (no counterpart in the
  // reference): per cell a mask, bit a a known face of axis a (fluid on one
  // side, or set by a layer), bit 3 + a a face with air on both sides; each of
  // P.extrap layers gives every such face not yet known the mean of the known
  // faces of the same axis in the six cells around (inside the tile grid, in
  // tiles with storage, summed x-, x+, y-, y+, z-, z+), new and old values alike.
  void extrapolate() {
    if (!P.extrap) return;
    const size_t C = size_t(scap) * 512u;
    std::vector<uint8_t> m0(C, 0), m1(C, 0);
    forActive([&](uint32_t c, I3 at) {
      const uint32_t th = ctypeOf(at.x, at.y, at.z);
      uint32_t m = 0;
      for (uint32_t axis = 0; axis < 3; axis++) {
        I3 lo = at;
        if (axis == 0) lo.x -= 1; else if (axis == 1) lo.y -= 1; else lo.z -= 1;
        const uint32_t tl = ctypeOf(lo.x, lo.y, lo.z);
        if (tl == 1u || th == 1u) m |= 1u << axis;
        else if (tl == 0u && th == 0u) m |= 8u << axis;
      }
      m0[c] = uint8_t(m);
    });
    const int32_t gx = int32_t(tx) * 8, gy = int32_t(ty) * 8, gz = int32_t(tz) * 8;
    for (uint32_t l = 0; l < P.extrap; l++) {
      const std::vector<uint8_t>& mi = (l & 1u) ? m1 : m0;
      std::vector<uint8_t>& mo = (l & 1u) ? m0 : m1;
      forActive([&](uint32_t c, I3 at) {
        const uint32_t m = mi[c];
        uint32_t o = m;
        for (uint32_t axis = 0; axis < 3; axis++) {
          if (((m >> axis) & 1u) || !((m >> (3u + axis)) & 1u)) continue;
          R sn = R(0), so = R(0);
          uint32_t cnt = 0;
          for (uint32_t q = 0; q < 6u; q++) {
            I3 r = at;
            const int32_t d = (q & 1u) ? 1 : -1;
            if (q < 2u) r.x += d; else if (q < 4u) r.y += d; else r.z += d;
            if (r.x < 0 || r.y < 0 || r.z < 0 || r.x >= gx || r.y >= gy || r.z >= gz) continue;
            const uint32_t sl = tmap[tileOf(r.x, r.y, r.z)];
            if (sl == NONE) continue;
            const uint32_t cn = sl * 512u + uint32_t(r.x & 7) + 8u * (uint32_t(r.y & 7) + 8u * uint32_t(r.z & 7));
            if (!((mi[cn] >> axis) & 1u)) continue;
            sn = sn + fvR(cn * 3u + axis);
            so = so + foR(cn * 3u + axis);
            cnt++;
          }
          if (cnt) {
            fvW(c * 3u + axis, sn / R(cnt));
            foW(c * 3u + axis, so / R(cnt));
            o |= 1u << axis;
          }
        }
        mo[c] = uint8_t(o);
      });
    }
  }
  void saveVel() {  // wgsl-narrowband.js:88-91
    forActive([&](uint32_t c, I3) { for (uint32_t ax = 0; ax < 3; ax++) velG[c * 3 + ax] = fvR(c * 3 + ax); });
  }
  V3<R> collideP(V3<R> pin) const {  // wgsl-solver.js:185-197
    V3<R> p = pin;
    R m = P.margin;
    for (uint32_t k = 0; k < P.ncols; k++) {
      if (!colNear(k, p)) continue;
      R phi; V3<R> wn;
      sdfEval(k, p, phi, wn);
      if (phi < P.q0) p = p + wn * (P.q0 - phi);
    }
    R big = R(1.0e9);
    V3<R> lo = v3((P.open & 1u) ? -big : P.x0 + m, (P.open & 4u) ? -big : P.y0 + m, (P.open & 16u) ? -big : P.z0 + m);
    V3<R> hi = v3((P.open & 2u) ? big : P.x1 - m, (P.open & 8u) ? big : P.y1 - m, (P.open & 32u) ? big : P.z1 - m);
    return v3(clampr(p.x, lo.x, hi.x), clampr(p.y, lo.y, hi.y), clampr(p.z, lo.z, hi.z));
  }
  static R fr(R x) { return x - M<R>::flo(x); }
  static R hsh3(V3<R> p) {  // wgsl-solver.js:198-201
    V3<R> q = v3(fr(p.x * R(0.3183099) + R(0.1)), fr(p.y * R(0.3183099) + R(0.2)), fr(p.z * R(0.3183099) + R(0.3)));
    q = q * R(17.0);
    return fr(q.x * q.y * q.z * (q.x + q.y + q.z));
  }
  static R vnz(V3<R> x) {  // 202-207
    V3<R> i = v3(M<R>::flo(x.x), M<R>::flo(x.y), M<R>::flo(x.z));
    V3<R> f = v3(fr(x.x), fr(x.y), fr(x.z));
    V3<R> u = v3(f.x * f.x * (R(3.0) - R(2.0) * f.x), f.y * f.y * (R(3.0) - R(2.0) * f.y), f.z * f.z * (R(3.0) - R(2.0) * f.z));
    R a = mixs(mixs(hsh3(i), hsh3(i + v3(R(1), R(0), R(0))), u.x), mixs(hsh3(i + v3(R(0), R(1), R(0))), hsh3(i + v3(R(1), R(1), R(0))), u.x), u.y);
    R b = mixs(mixs(hsh3(i + v3(R(0), R(0), R(1))), hsh3(i + v3(R(1), R(0), R(1))), u.x), mixs(hsh3(i + v3(R(0), R(1), R(1))), hsh3(i + v3(R(1), R(1), R(1))), u.x), u.y);
    return mixs(a, b, u.z);
  }
  void stageG2P() {  // g2p 208-275
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < int64_t(P.n); ii++) {
      size_t i = size_t(ii);
      if (pos[i * 4 + 1] < R(-1.0e5)) continue;
      V3<R> p = v3(pos[i * 4], pos[i * 4 + 1], pos[i * 4 + 2]);
      V3<R> v0 = v3(vel[i * 4], vel[i * 4 + 1], vel[i * 4 + 2]);
      I3 fcell = particleCell(p);
      V3<R> vn = v0 - v3(R(0), P.gravity * P.dt, R(0));
      V3<R> vo = v0;
      if (hasStorage(fcell)) sampleBoth(p, vn, vo);
      V3<R> fl = v0 + (vn - vo);
      V3<R> v = v3(mixs(vn.x, fl.x, P.flip), mixs(vn.y, fl.y, P.flip), mixs(vn.z, fl.z, P.flip));
      if (P.ocean != 0u) {  // wgsl-solver.js:218
        R lw = layerW(p);
        if (lw > R(0)) {
          R sc = clampr(lw * R(2), R(0), R(1));
          V3<R> ov = oceanVel(p);
          R t = sc * sc * (R(1) - M<R>::exp(R(-8.0) * P.dt));
          v = v3(mixs(v.x, ov.x, t), mixs(v.y, ov.y, t), mixs(v.z, ov.z, t));
        }
      }
      I3 bc{toi(M<R>::flo((p.x - P.x0) * P.ih - R(0.5))), toi(M<R>::flo((p.y - P.y0) * P.ih - R(0.5))), toi(M<R>::flo((p.z - P.z0) * P.ih - R(0.5)))};
      R csum = R(8.0) * P.rest;
      bool fast = P.q12 > R(0.5);
      if (P.q9 < R(0.5) && fast) {
        csum = R(8.0) * R(cnt[cellIdx(fcell.x, fcell.y, fcell.z)]);
      } else if (P.q9 < R(0.5)) {
        csum = R(0.0);
        for (int32_t dz = 0; dz <= 1; dz++)
          for (int32_t dy = 0; dy <= 1; dy++)
            for (int32_t dx = 0; dx <= 1; dx++) {
              I3 cc{iclamp(bc.x + dx, 0, int32_t(P.nx) - 1), iclamp(bc.y + dy, 0, int32_t(P.ny) - 1), iclamp(bc.z + dz, 0, int32_t(P.nz) - 1)};
              csum = csum + R(cnt[cellIdx(cc.x, cc.y, cc.z)]);
            }
      }
      R rho = csum / (R(8.0) * P.rest);
      R air = R(1.0) - smooth(R(0.008), R(0.1), rho);
      R shield = smooth(R(0.1), R(0.65), rho);
      R exposure = R(1.0) - R(0.97) * shield;
      R spd0 = len(v);
      R sens = fmx<R>(P.q6, R(0.05));
      R pacc = len(vn - vo) / P.dt;
      R settle = smooth(R(0.1), R(0.5), P.q8);
      R impact = clampr((pacc / sens - R(15.0)) / R(80.0), R(0.0), R(1.0)) * settle;
      R turb = clampr((len(v0 - vo) / sens - R(0.35)) / R(1.5), R(0.0), R(1.0)) * settle;
      R spray = air * (R(0.25) + R(0.75) * clampr(spd0 / R(4.0), R(0.0), R(1.0)));
      R tau = fmx<R>(P.q5, R(0.05)) * mixs(R(1.0), R(0.35), smooth(R(0.7), R(1.0), rho)) * mixs(R(1.0), R(2.5), air);
      R foam = vel[i * 4 + 3] * M<R>::exp(-P.dt / tau);
      foam = fmx<R>(foam, fmx<R>(spray * R(0.9), fmx<R>(impact, turb * R(0.8))));
      if (P.ocean != 0u) foam = foam * (R(1) - clampr(layerW(p) * R(2), R(0), R(1)));  // wgsl-solver.js:248
      V3<R> gust = v3(R(0), R(0), R(0));
      if (P.q7 > R(0.0) && P.q4 > R(0.0) && (!fast || shield < R(0.9))) {
        V3<R> gp = p * R(0.55);
        R tt = P.q8 * R(0.6);
        gust = v3(vnz(gp + v3(tt, R(0.0), R(0.0))), vnz(gp + v3(R(17.3), tt, R(4.1))), vnz(gp + v3(R(5.7), R(9.2), tt))) * R(2.0) - v3(R(1), R(1), R(1));
      }
      R kd = P.q4 * (R(0.04) + R(1.5) * exposure * exposure);
      V3<R> airv = v3(P.q1, P.q2, P.q3) + gust * P.q7;
      v = v + (airv - v) * (R(1.0) - M<R>::exp(-kd * P.dt));
      R sp = len(v);
      if (sp > P.vmax) v = v * (P.vmax / sp);
      V3<R> pt = p + v * P.dt;
      if (P.rk2 != 0u && hasStorage(fcell)) {
        // opt-in midpoint rule through the grid's new velocity (no counterpart in
        // the reference): pt = p + dt u(p + dt/2 u(p)), vn being u(p)
        const R hdt = R(0.5) * P.dt;
        V3<R> pm = v3(p.x + vn.x * hdt, p.y + vn.y * hdt, p.z + vn.z * hdt), um, uo;
        sampleBoth(pm, um, uo);
        pt = v3(p.x + um.x * P.dt, p.y + um.y * P.dt, p.z + um.z * P.dt);
      }
      V3<R> pc = collideP(pt);
      V3<R> dsp = pc - pt;
      if (dot(dsp, dsp) > R(1.0e-12)) {
        V3<R> nrm = dsp / len(dsp);
        R vd = dot(v, nrm);
        if (vd < R(0.0)) v = v - nrm * vd;
      }
      uint32_t o = P.open;
      if (o != 0u && (((o & 1u) && pc.x < P.x0) || ((o & 2u) && pc.x > P.x1) || ((o & 4u) && pc.y < P.y0) || ((o & 8u) && pc.y > P.y1) ||
                      ((o & 16u) && pc.z < P.z0) || ((o & 32u) && pc.z > P.z1))) {
        pos[i * 4] = R(0); pos[i * 4 + 1] = R(-1.0e6); pos[i * 4 + 2] = R(0); pos[i * 4 + 3] = R(0);
        vel[i * 4] = R(0); vel[i * 4 + 1] = R(0); vel[i * 4 + 2] = R(0); vel[i * 4 + 3] = R(0);
        continue;
      }
      pos[i * 4] = pc.x; pos[i * 4 + 1] = pc.y; pos[i * 4 + 2] = pc.z;
      pos[i * 4 + 3] = fmn<R>(len(v), R(63.0)) + R(64.0) * M<R>::flo(foam * R(15.0) + R(0.5));
      vel[i * 4] = v.x; vel[i * 4 + 1] = v.y; vel[i * 4 + 2] = v.z; vel[i * 4 + 3] = foam;
    }
  }
  // ---- narrow band (wgsl-narrowband.js) ----
  R jacNb(I3 p, R cur, bool fromRhs) const {  // 51-64
    if (!inDom(p)) return R(8.0);
    if (cur >= R(0.0)) return cur;
    R m = R(-8.0);
    for (uint32_t q = 0; q < 6; q++) {
      I3 n = p;
      if (q == 0) n.x -= 1; else if (q == 1) n.x += 1; else if (q == 2) n.y -= 1; else if (q == 3) n.y += 1; else if (q == 4) n.z -= 1; else n.z += 1;
      if (!inDom(n)) continue;
      uint32_t cq = cellIdx(n.x, n.y, n.z);
      if (colBits(cq) != 0u) continue;
      m = fmx<R>(m, fromRhs ? rhs[cq] : phiOf(cq));
    }
    return clampr(m - R(1.0), R(-7.9), R(-0.5));
  }
  void nbPhi() {  // 39-49
    forActive([&](uint32_t c, I3 p) {
      if (!inDom(p)) { rhs[c] = R(8.0); return; }
      if (colBits(c) != 0u) { rhs[c] = R(0.5); return; }
      V3<R> x = cellCenter(p.x, p.y, p.z);
      R phA = phiSample(x - velGVec(x) * P.dt);
      bool inside = cnt[c] > 0 || phA < R(-1.0);
      if (P.ocean != 0u && layerW(x) >= R(0.5)) inside = x.y < oceanEta(x.x, x.z);
      rhs[c] = inside ? clampr(phA, R(-7.9), R(-0.5)) : clampr(phA, R(0.5), R(7.9));
    });
  }
  void nbJacA() { forActive([&](uint32_t c, I3 p) { phiPut(c, jacNb(p, rhs[c], true)); }); }   // 65-66
  void nbJacB() { forActive([&](uint32_t c, I3 p) { rhs[c] = jacNb(p, phiOf(c), false); }); }  // 67-68
  void stageNarrowBand() {  // the 'narrow band' stage (sim.js:281-283)
    nbPhi();
    nbJacA();
    nbJacB();
    nbJacA();
  }
  static R hash1(uint32_t n0) {  // wgsl-solver.js:181-184
    uint32_t x = n0;
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return R(float(x)) * R(2.3283064e-10);
  }
  // ---- whitewater and the air field (wgsl-whitewater.js:1-163) ----
  R wetCell(int32_t i, int32_t j, int32_t k) const {  // 7-13
    if (!inDom(I3{i, j, k}) || tmap[tileOf(i, j, k)] == NONE) return R(0);
    uint32_t c = cellIdx(i, j, k);
    if (P.nb != 0u && (ctype[c] & 0xffu) == 1u && phiOf(c) < R(-1.0)) return R(1);
    return fmn(R(cnt[c]) / P.rest, R(1));
  }
  R wetAt(V3<R> p) const {  // 14-22
    V3<R> g = v3((p.x - P.x0) * P.ih - R(0.5), (p.y - P.y0) * P.ih - R(0.5), (p.z - P.z0) * P.ih - R(0.5));
    V3<R> f0 = v3(M<R>::flo(g.x), M<R>::flo(g.y), M<R>::flo(g.z));
    V3<R> fr = g - f0;
    I3 i0{toi(f0.x), toi(f0.y), toi(f0.z)};
    R s = R(0);
    for (uint32_t c = 0; c < 8; c++) {
      int32_t dx = int32_t(c & 1u), dy = int32_t((c >> 1) & 1u), dz = int32_t((c >> 2) & 1u);
      R w = (dx == 1 ? fr.x : R(1) - fr.x) * (dy == 1 ? fr.y : R(1) - fr.y) * (dz == 1 ? fr.z : R(1) - fr.z);
      if (w > R(0)) s = s + w * wetCell(i0.x + dx, i0.y + dy, i0.z + dz);
    }
    return s;
  }
  uint32_t airIdx(int32_t i, int32_t j, int32_t k) const {  // 24-27
    i = iclamp(i, 0, int32_t(P.aNx) - 1);
    j = iclamp(j, 0, int32_t(P.aNy) - 1);
    k = iclamp(k, 0, int32_t(P.aNz) - 1);
    return uint32_t(i) + P.aNx * (uint32_t(j) + P.aNy * uint32_t(k));
  }
  V3<R> airCellV(uint32_t c) const { return v3(air[c * 4], air[c * 4 + 1], air[c * 4 + 2]); }
  V3<R> airAt(V3<R> p) const {  // 29-39
    V3<R> g = v3((p.x - P.x0) / P.airH - R(0.5), (p.y - P.y0) / P.airH - R(0.5), (p.z - P.z0) / P.airH - R(0.5));
    V3<R> f0 = v3(M<R>::flo(g.x), M<R>::flo(g.y), M<R>::flo(g.z));
    V3<R> fr = g - f0;
    I3 i0{toi(f0.x), toi(f0.y), toi(f0.z)};
    V3<R> s = v3(R(0), R(0), R(0));
    for (uint32_t c = 0; c < 8; c++) {
      int32_t dx = int32_t(c & 1u), dy = int32_t((c >> 1) & 1u), dz = int32_t((c >> 2) & 1u);
      R w = (dx == 1 ? fr.x : R(1) - fr.x) * (dy == 1 ? fr.y : R(1) - fr.y) * (dz == 1 ? fr.z : R(1) - fr.z);
      s = s + airCellV(airIdx(i0.x + dx, i0.y + dy, i0.z + dz)) * w;
    }
    return s;
  }
  void airPush(V3<R> p, V3<R> v, R w) {  // 40-45
    uint32_t c = airIdx(toi(M<R>::flo((p.x - P.x0) / P.airH)), toi(M<R>::flo((p.y - P.y0) / P.airH)), toi(M<R>::flo((p.z - P.z0) / P.airH))) * 4u;
    airAcc[c] += toi(v.x * w * R(256));
    airAcc[c + 1] += toi(v.y * w * R(256));
    airAcc[c + 2] += toi(v.z * w * R(256));
    airAcc[c + 3] += toi(w * R(256));
  }
  void stageSpray() {  // the 'spray' stage (sim.js:296-305)
    const uint32_t used = std::min(sctl[0], P.smax);
    if (airOnStep) {
      std::fill(airAcc.begin(), airAcc.end(), 0);  // airClear
      for (uint32_t i = 0; i < P.n; i++) {  // airSplatW 54-59
        if (pos[i * 4 + 1] < R(-1.0e5)) continue;
        V3<R> p = v3(pos[i * 4], pos[i * 4 + 1], pos[i * 4 + 2]);
        I3 c = particleCell(p);
        if (hasStorage(c) && R(cnt[cellIdx(c.x, c.y, c.z)]) > R(0.85) * P.rest) continue;
        airPush(p, v3(vel[i * 4], vel[i * 4 + 1], vel[i * 4 + 2]), R(1));
      }
      for (uint32_t j = 0; j < used; j++) {  // airSplatS 61-66
        if (sPos[j * 4 + 3] <= R(0) || sVel[j * 4 + 3] > R(1.5)) continue;
        airPush(v3(sPos[j * 4], sPos[j * 4 + 1], sPos[j * 4 + 2]), v3(sVel[j * 4], sVel[j * 4 + 1], sVel[j * 4 + 2]), R(0.25));
      }
      for (uint32_t c = 0; c < airN; c++) {  // airUpdate 68-86
        int32_t i = int32_t(c % P.aNx), j = int32_t((c / P.aNx) % P.aNy), k = int32_t(c / (P.aNx * P.aNy));
        V3<R> a = airCellV(c);
        V3<R> nb = airCellV(airIdx(i + 1, j, k)) + airCellV(airIdx(i - 1, j, k)) + airCellV(airIdx(i, j + 1, k)) +
                   airCellV(airIdx(i, j - 1, k)) + airCellV(airIdx(i, j, k + 1)) + airCellV(airIdx(i, j, k - 1));
        nb = nb / R(6);
        a = v3(mixs(a.x, nb.x, P.airBlur), mixs(a.y, nb.y, P.airBlur), mixs(a.z, nb.z, P.airBlur));
        R w = R(airAcc[c * 4 + 3]) / R(256);
        if (w > R(0)) {
          V3<R> vt = v3(R(airAcc[c * 4]), R(airAcc[c * 4 + 1]), R(airAcc[c * 4 + 2])) / (w * R(256));
          R t = R(1) - M<R>::exp(-P.airCouple * fmn(w, R(8)) * P.dt);
          a = v3(mixs(a.x, vt.x, t), mixs(a.y, vt.y, t), mixs(a.z, vt.z, t));
        }
        V3<R> x = v3(P.x0 + (R(i) + R(0.5)) * P.airH, P.y0 + (R(j) + R(0.5)) * P.airH, P.z0 + (R(k) + R(0.5)) * P.airH);
        for (uint32_t q = 0; q < P.ncols; q++) {
          if (!colNear(q, x)) continue;
          R phi;
          V3<R> wn;
          sdfEval(q, x, phi, wn);
          if (phi < R(0.6) * P.airH) { a = colSurfVel(q, x); break; }
        }
        R t = R(1) - M<R>::exp(-P.airDecay * P.dt);
        a = v3(mixs(a.x, P.q1, t), mixs(a.y, P.q2, t), mixs(a.z, P.q3, t));
        air2[c * 4] = a.x; air2[c * 4 + 1] = a.y; air2[c * 4 + 2] = a.z; air2[c * 4 + 3] = R(0);
      }
      air = air2;  // airCopy
    }
    for (uint32_t j = 0; j < used; j++) {  // wwSample 88-95
      if (sPos[j * 4 + 3] <= R(0)) continue;
      V3<R> p = v3(sPos[j * 4], sPos[j * 4 + 1], sPos[j * 4 + 2]), vf = v3(R(0), R(0), R(0)), vo;
      if (hasStorage(particleCell(p))) sampleBoth(p, vf, vo);
      sAux[j * 4] = vf.x; sAux[j * 4 + 1] = vf.y; sAux[j * 4 + 2] = vf.z; sAux[j * 4 + 3] = wetAt(p);
    }
    const R hold = R(0.2f);  // WWHOLD
    for (uint32_t j = 0; j < used; j++) {  // sprayStep 129-181
      if (sPos[j * 4 + 3] <= R(0)) continue;
      V3<R> v = v3(sVel[j * 4], sVel[j * 4 + 1], sVel[j * 4 + 2]);
      R kw = sVel[j * 4 + 3];
      uint32_t kind = tou(kw + R(0.5));
      R tk = (kw - R(kind)) / R(0.4) * hold;
      const bool settled = tk >= hold * R(0.999f);
      const uint32_t k0 = kind;
      R wet = sAux[j * 4 + 3];
      V3<R> vf = v3(sAux[j * 4], sAux[j * 4 + 1], sAux[j * 4 + 2]);
      R life = sPos[j * 4 + 3];
      uint32_t h0 = j * 2654435761u + P.seed * 974711u;
      if (kind == 0u && v.y < vf.y && wet > R(0.55) && tk > R(0.05)) { kind = 2u; life = P.wwFoamLife * (R(0.6) + R(0.8) * hash1(h0)); v = vf; }
      else if (kind == 1u && wet > R(0.55)) { life = R(0); }
      else if (kind == 2u && wet > R(0.97) && settled) { kind = 3u; }
      else if (kind == 2u && wet < R(0.04) && settled) { kind = 0u; life = fmn(life, P.sprLife); }
      else if (kind == 3u && wet < R(0.6) && settled) { kind = 2u; life = fmx(life, R(0.5) * P.wwFoamLife); }
      if (kind != k0) tk = R(0); else tk = fmn(tk + P.dt, hold);
      V3<R> p = v3(sPos[j * 4], sPos[j * 4 + 1], sPos[j * 4 + 2]);
      if (kind <= 1u) {
        const bool mist = kind == 1u;
        V3<R> gust = v3(R(0), R(0), R(0));
        if (P.q7 > R(0)) {
          V3<R> gp = p * R(0.55) * P.vs;
          R tt = P.q8 * R(0.6);
          gust = v3(vnz(gp + v3(tt, R(0), R(0))), vnz(gp + v3(R(17.3), tt, R(4.1))), vnz(gp + v3(R(5.7), R(9.2), tt))) * R(2) - v3(R(1), R(1), R(1));
        }
        V3<R> base = v3(P.q1, P.q2, P.q3);
        if (P.airOn != 0u) base = airAt(p);
        V3<R> airv = base + gust * P.q7 * (mist ? R(2.5) : R(1));
        v.y = v.y - P.gravity * P.dt;
        v = v + (airv - v) * (R(1) - M<R>::exp(-(mist ? P.mistDrag : P.sprDrag) * P.dt));
        life = life - P.dt;
      } else if (kind == 2u) {
        R vy = vf.y + clampr((wet - R(0.5)) * R(4), R(-1), R(1)) * P.h * R(4);
        R t = R(1) - M<R>::exp(R(-10) * P.dt);
        v = v3(mixs(v.x, vf.x, t), mixs(v.y, vy, t), mixs(v.z, vf.z, t));
        life = life - P.dt;
      } else {
        V3<R> tv = vf + v3(R(0), P.wwRise, R(0));
        R t = R(1) - M<R>::exp(R(-6) * P.dt);
        v = v3(mixs(v.x, tv.x, t), mixs(v.y, tv.y, t), mixs(v.z, tv.z, t));
        life = life - R(0.3) * P.dt;
      }
      V3<R> np = p + v * P.dt;
      if (np.x < P.x0 || np.x > P.x1 || np.y < P.y0 || np.y > P.y1 || np.z < P.z0 || np.z > P.z1) life = R(0);
      else if (P.ncols > 0u && colAt(np) >= 0) life = R(0);
      else if (P.ocean != 0u && kind >= 2u && layerW(np) > R(0.6)) life = R(0);
      sPos[j * 4] = np.x; sPos[j * 4 + 1] = np.y; sPos[j * 4 + 2] = np.z; sPos[j * 4 + 3] = life;
      sVel[j * 4] = v.x; sVel[j * 4 + 1] = v.y; sVel[j * 4 + 2] = v.z; sVel[j * 4 + 3] = R(kind) + R(0.4) * tk / hold;
    }
    if (sprOnStep) sprayEmit();
    if (dripOnStep) stageDrips();
  }
  void sprayEmit() {
    uint32_t e = 0;  // sprayEmit 98-127: points in particle order, the ring's next slots
    const uint32_t first = sctl[0];
    for (uint32_t i = 0; i < P.n; i++) {
      if (pos[i * 4 + 1] < R(-1.0e5)) continue;
      V3<R> q = v3(pos[i * 4], pos[i * 4 + 1], pos[i * 4 + 2]), v = v3(vel[i * 4], vel[i * 4 + 1], vel[i * 4 + 2]);
      if (P.ocean != 0u && layerW(q) > R(0)) continue;  // wgsl-whitewater.js:105
      R foam = vel[i * 4 + 3], spd = len(v);
      R fast = clampr((spd * P.vs - P.sprMin) / (R(2) * P.sprMin), R(0), R(1));
      I3 bc{toi(M<R>::flo((q.x - P.x0) * P.ih - R(0.5))), toi(M<R>::flo((q.y - P.y0) * P.ih - R(0.5))), toi(M<R>::flo((q.z - P.z0) * P.ih - R(0.5)))};
      R rho = R(0);
      for (int32_t dz = 0; dz <= 1; dz++)
        for (int32_t dy = 0; dy <= 1; dy++)
          for (int32_t dx = 0; dx <= 1; dx++) rho = rho + wetCell(bc.x + dx, bc.y + dy, bc.z + dz);
      rho = rho / R(8);
      R crest = (R(1) - smooth(R(0.3), R(0.75), rho)) * fast * (R(0.3) + R(0.7) * foam);
      R churn = (R(1) - smooth(R(0.8), R(1), rho)) * clampr((foam - R(0.55)) / R(0.45), R(0), R(1)) * smooth(R(0.5), R(1.5), spd * P.vs);
      R em = (P.wwCrest * crest + P.wwImpact * churn) * smooth(R(0.1), R(0.5), P.q8);
      if (em <= R(0)) continue;
      uint32_t h0 = i * 747796405u + P.seed * 2891336453u;
      uint32_t k = std::min(tou(M<R>::flo(P.sprRate * em * P.dt + hash1(h0))), 4u);
      for (uint32_t m = 0; m < k; m++, e++) {
        if (e >= P.smax) continue;  // the ring came round to points placed this substep
        uint32_t hh = h0 ^ ((m + 1u) * 0x9E3779B9u);
        R z = hash1(hh + 1u) * R(2) - R(1), a = hash1(hh + 2u) * R(6.2831853f), sr = M<R>::sq(fmx(R(0), R(1) - z * z));
        R sa, ca;
        M<R>::sincos(a, &sa, &ca);
        V3<R> dir = v3(sr * ca, sr * sa, z);
        bool under = rho > R(0.7) && hash1(hh + 7u) < R(0.7);
        bool mist = !under && hash1(hh + 3u) < P.mistFrac * fast;
        V3<R> nv = v * (R(0.9) + R(0.2) * hash1(hh + 4u)) + dir * spd * (mist ? R(0.2) : R(0.06));
        R life = P.sprLife * (R(0.6) + R(0.8) * hash1(hh + 5u)) * (mist ? R(1.6) : R(1));
        R kind = mist ? R(1) : R(0);
        if (under) { nv = v + dir * spd * R(0.05); life = P.wwFoamLife * (R(0.6) + R(0.8) * hash1(hh + 5u)); kind = R(3); }
        uint32_t slot = (first + e) % P.smax;
        if (sPos[slot * 4 + 3] > R(0)) continue;
        V3<R> np = q + nv * P.dt + dir * P.h * R(0.3) * hash1(hh + 6u);
        sPos[slot * 4] = np.x; sPos[slot * 4 + 1] = np.y; sPos[slot * 4 + 2] = np.z; sPos[slot * 4 + 3] = life;
        sVel[slot * 4] = nv.x; sVel[slot * 4 + 1] = nv.y; sVel[slot * 4 + 2] = nv.z; sVel[slot * 4 + 3] = kind;
      }
    }
    sctl[0] = first + e;
  }
  void stageDrips() {  // dripStep (wgsl-whitewater.js:256-282), samples in order, drops in the ring's next slots
    const uint32_t first = sctl[0];
    uint32_t e = 0;
    for (uint32_t sIdx = 0; sIdx < P.ndrip; sIdx++) {
      R* a = &drip[size_t(sIdx) * 8];
      R* b = a + 4;
      uint32_t k = tou(a[3]);
      if (k >= P.ncols) continue;
      V3<R> r0 = colRow(k, 2), r1 = colRow(k, 3), r2 = colRow(k, 4), c = colRow(k, 5);
      V3<R> wp = c + r0 * a[0] + r1 * a[1] + r2 * a[2];
      V3<R> n = r0 * b[0] + r1 * b[1] + r2 * b[2];
      V3<R> wn = n / len(n);
      V3<R> probe = wp + wn * P.h * R(0.75);
      R here = wetAt(probe);
      if (here > R(0.4)) { b[3] = R(1); continue; }
      R wetness = fmx(R(0), b[3] - P.dt / fmx(P.dripDry, R(0.05)));
      b[3] = wetness;
      if (wetness <= R(0) || here > R(0.1) || P.smax == 0u || P.dripRate <= R(0)) continue;
      R down = clampr(-wn.y, R(0), R(1));
      R rate = P.dripRate * wetness * wetness * (R(0.15) + R(0.85) * down);
      uint32_t h0 = sIdx * 747796405u + P.seed * 2891336453u + 17u;
      if (hash1(h0) >= rate * P.dt) continue;
      uint32_t eh = e++;
      if (eh >= P.smax) continue;
      uint32_t slot = (first + eh) % P.smax;
      if (sPos[slot * 4 + 3] > R(0)) continue;
      V3<R> v = colSurfVel(k, wp) + wn * P.h * R(2) * hash1(h0 + 1u);
      sPos[slot * 4] = probe.x; sPos[slot * 4 + 1] = probe.y; sPos[slot * 4 + 2] = probe.z; sPos[slot * 4 + 3] = P.sprLife * R(2);
      sVel[slot * 4] = v.x; sVel[slot * 4 + 1] = v.y; sVel[slot * 4 + 2] = v.z; sVel[slot * 4 + 3] = R(0);
    }
    sctl[0] = first + e;
  }
  // deepClear/deepPick/deepFin (wgsl-whitewater.js:279-297): the centre of every fluid
  // cell in the domain holding no particle, w = 4096 + min(|velG there|, 63); above
  // 200000 such cells at the previous read, the cells with hash1(c 2654435761 + 7)
  // prev < 200000.  In active-cell order (slot list, then cell), the first 200000.
  uint32_t deepPrev = 0;
  uint32_t readDeep(double* out, uint32_t room) {
    const uint32_t DEEPMAX = 200000u;
    if (!nbOn || !velGOn || !P.n) return 0;
    uint32_t all = 0, k = 0;
    for (uint32_t b = 0; b < ctl[4]; b++) {
      const uint32_t s = slotList[b];
      for (uint32_t local = 0; local < 512u; local++) {
        const uint32_t c = s * 512u + local;
        const I3 p = slotCoord(s, local);
        if (!inDom(p) || (ctype[c] & 0xffu) != 1u || cnt[c] > 0) continue;
        all++;
        if (deepPrev > DEEPMAX && float(hash1(c * 2654435761u + 7u)) * float(deepPrev) >= float(DEEPMAX)) continue;
        if (k >= DEEPMAX) continue;
        const V3<R> x = cellCenter(p.x, p.y, p.z);
        if (k < room) {
          out[k * 4 + 0] = double(x.x);
          out[k * 4 + 1] = double(x.y);
          out[k * 4 + 2] = double(x.z);
          out[k * 4 + 3] = double(R(4096.0) + std::min(len(velGVec(x)), R(63.0)));
        }
        k++;
      }
    }
    deepPrev = all;
    return std::min(k, room);
  }
  void setDrips(const float* a, uint32_t n) {
    drip.assign(size_t(n) * 8, R(0));
    for (size_t q = 0; q < size_t(n) * 8; q++) drip[q] = R(a[q]);
    ndrip = n;
  }
  void setSprayCap(uint32_t n) {  // sim.js:333-340
    if (n == sprayCap) return;
    sprayCap = n;
    sPos.assign(size_t(n) * 4, R(0));
    sVel.assign(size_t(n) * 4, R(0));
    sAux.assign(size_t(n) * 4, R(0));
  }
  void nbKill() {  // wgsl-narrowband.js:93-104
    for (uint32_t i = 0; i < P.n; i++) {
      if (pos[i * 4 + 1] < R(-1.0e5)) continue;
      V3<R> q = v3(pos[i * 4], pos[i * 4 + 1], pos[i * 4 + 2]);
      R lw = layerW(q);
      bool kill = lw >= R(0.5) || (lw > R(0.25) && q.y > oceanEta(q.x, q.z) + R(0.5) * P.h);
      V3<R> v = v3(vel[i * 4], vel[i * 4 + 1], vel[i * 4 + 2]);
      if (!kill && P.nb != 0u && calmV(v)) {
        I3 cc = particleCell(v3(pos[i * 4], pos[i * 4 + 1], pos[i * 4 + 2]));
        if (hasStorage(cc) && phiOf(cellIdx(cc.x, cc.y, cc.z)) < -P.band) kill = true;
      }
      if (kill) {
        pos[i * 4] = R(0); pos[i * 4 + 1] = R(-1.0e6); pos[i * 4 + 2] = R(0); pos[i * 4 + 3] = R(0);
        vel[i * 4] = R(0); vel[i * 4 + 1] = R(0); vel[i * 4 + 2] = R(0); vel[i * 4 + 3] = R(0);
      }
    }
  }
  void stageReseed() {  // nbKill .. nbSpawn (wgsl-narrowband.js:93-149)
    nbKill();
    ctl[7] = ctl[8] = ctl[9] = ctl[10] = 0;  // nbReset
    for (uint32_t i = 0; i < P.n; i++) {      // nbDead, in index order
      if (pos[i * 4 + 1] < R(-1.0e5)) { uint32_t k = ctl[8]++; if (k < lcap) lst[k] = i; }
    }
    uint32_t sc = ctl[4];                     // nbDeficit, in slotList order
    for (uint32_t b = 0; b < sc; b++) {
      uint32_t s = slotList[b];
      for (uint32_t local = 0; local < 512u; local++) {
        uint32_t c = s * 512u + local;
        I3 p = slotCoord(s, local);
        if (!inDom(p) || colBits(c) != 0u) continue;
        V3<R> cx = cellCenter(p.x, p.y, p.z);
        R lw = layerW(cx);
        R ph = phiOf(c);
        if (lw > R(0.25) && lw < R(0.5)) {
          if (cx.y > oceanEta(cx.x, cx.z) - R(0.25) * P.h || ph < -P.band) continue;
        } else {
          if ((ctype[c] & 0xffu) != 1u || ph > R(-1.0) || ph < -P.band) continue;
          if (!calmV(velGVec(cx))) continue;
        }
        uint32_t n = uint32_t(imax(cnt[c], 0));
        if (R(n) >= R(0.55) * P.rest) continue;
        uint32_t needc = std::min(tou(M<R>::rnd(R(0.85) * P.rest)) - n, 64u);
        uint32_t k = ctl[9]++;
        if (k >= lcap / 2u) continue;
        lst[lcap + k * 2] = uint32_t(p.x) | (uint32_t(p.y) << 16);
        lst[lcap + k * 2 + 1] = uint32_t(p.z) | (needc << 16);
      }
    }
    uint32_t ne = std::min(ctl[9], lcap / 2u);  // nbSpawn, entries in order
    for (uint32_t e = 0; e < ne; e++) {
      uint32_t a0 = lst[lcap + e * 2], a1 = lst[lcap + e * 2 + 1];
      V3<R> cp = v3(R(a0 & 0xffffu), R(a0 >> 16), R(a1 & 0xffffu));
      uint32_t needc = a1 >> 16;
      uint32_t base = ctl[10];
      ctl[10] += needc;
      uint32_t avail = std::min(ctl[8], lcap);
      for (uint32_t m = 0; m < needc; m++) {
        uint32_t k = base + m;
        if (k >= avail) break;
        uint32_t i = lst[k];
        uint32_t hh = (e * 9781u + m * 6271u + 17u) ^ (P.seed * 2654435761u);
        V3<R> x = v3(P.x0, P.y0, P.z0) + (cp + v3(hash1(hh + 1u), hash1(hh + 2u), hash1(hh + 3u))) * P.h;
        V3<R> v = velGVec(x);
        if (layerW(x) > R(0.25)) v = oceanVel(x);
        pos[i * 4] = x.x; pos[i * 4 + 1] = x.y; pos[i * 4 + 2] = x.z; pos[i * 4 + 3] = fmn<R>(len(v), R(63.0));
        vel[i * 4] = v.x; vel[i * 4 + 1] = v.y; vel[i * 4 + 2] = v.z; vel[i * 4 + 3] = R(0);
      }
    }
  }

  // ---- host side of FlipSim.step (sim.js:179-241) ----
  int beginStep(const st_params* o) {
    if (N == 0) return ST_ERR_STATE;
    nbOn = o->nb_on != 0 || o->ocean_on != 0;
    if (nbOn && !velGOn) {  // enableNB (sim.js:120-126): longer lists
      velGOn = true;
      lcap = scap * 512u;
      lst.assign(size_t(lcap) * 2, 0u);
    }
    initStep = nbOn ? o->nb_init : 0u;
    uint32_t nfv = initStep ? o->nfv : 0u;
    uint32_t nc = ncolsTotal >= o->nfv ? ncolsTotal - o->nfv : 0u;
    // solid cache (sim.js:191-195)
    {
      bool same = haveCols && sdfSig == sdfSigPrev && !o->no_solid_cache && cfPrev.size() == size_t(nc) * 32;
      if (same)
        for (size_t q = 0; q < size_t(nc) * 32; q++)
          if (cfPrev[q] != float(cols[q])) { same = false; break; }
      if (!same) {
        solidDirty = true;
        cfPrev.resize(size_t(nc) * 32);
        for (size_t q = 0; q < size_t(nc) * 32; q++) cfPrev[q] = float(cols[q]);
        sdfSigPrev = sdfSig;
        haveCols = true;
      }
    }
    onlyTiles = o->only_tiles != 0;
    Params<R>& p = P;
    p.n = N; p.nx = uint32_t(nx); p.ny = uint32_t(ny); p.nz = uint32_t(nz); p.ncols = nc; p.open = open; p.ncell = ncell;
    double hd = hd_;
    double omegaDef = std::min(1.9, 2.0 / (1.0 + std::sin(M_PI / std::max(nx, std::max(ny, nz)))));
    p.dt = R(o->dt); p.h = h; p.ih = R(1.0 / hd); p.gravity = R(o->g); p.flip = R(o->flip); p.rest = rest; p.comp = R(o->comp);
    p.omega = o->omega != 0.0f ? R(o->omega) : R(omegaDef);
    p.x0 = bx0; p.y0 = by0; p.z0 = bz0; p.vmax = o->vmax != 0.0f ? R(o->vmax) : R(14);
    p.margin = R(0.3 * dd_); p.scale = R(65536); p.x1 = bx1; p.y1 = by1; p.z1 = bz1; p.q0 = R(0.3 * hd);
    p.q1 = R(o->wind[0]); p.q2 = R(o->wind[1]); p.q3 = R(o->wind[2]); p.q4 = R(o->drag); p.q5 = R(o->foam_tau); p.q6 = R(o->foam_sens);
    p.q7 = R(o->gust); p.q8 = R(tm); p.q9 = o->no_foam ? R(1) : R(0); p.q10 = o->no_cull ? R(1) : R(0); p.q11 = R(1); p.q12 = o->fast_air ? R(1) : R(0);
    if (!onlyTiles) tm = tm + double(o->dt);
    p.lcap = lcap; p.tx = tx; p.ty = ty; p.tz = tz; p.scap = scap;
    p.sheetMin = R(o->tear);
    if (!onlyTiles) seed = seed + 1u;
    p.seed = seed;
    p.nb = nbOn ? 1u : 0u;
    p.band = R(std::max(2.0f, o->nb_band > 0 ? o->nb_band : 3.0f));
    p.ocean = o->ocean_on ? 1u : 0u;
    p.rk2 = o->rk2 ? 1u : 0u;
    p.extrap = o->extrap_layers;
    p.nwave = o->ocean_on ? std::min(o->nwave, 32u) : 0u;
    p.otime = R(o->ocean_time);
    p.olevel = o->ocean_on ? R(o->ocean_level) : R(0);
    p.olayer = o->ocean_on ? R(o->ocean_layer) : R(0);
    p.oy0 = by0;
    p.odepth = o->ocean_on ? R(float(std::max(double(o->ocean_level) - double(float(by0)), 1e-3))) : R(1);
    for (uint32_t k = 0; k < p.nwave; k++)
      for (int q = 0; q < 8; q++) p.waves[k * 2 + q / 4][q % 4] = R(o->waves[k * 8 + q]);
    p.nbInit = initStep;
    p.nfill = initStep ? std::min<uint32_t>(o->nfill, 8u) : 0u;
    p.nfv = nfv;
    for (int k = 0; k < 16; k++) for (int q = 0; q < 4; q++) p.fills[k][q] = R(o->fills[k * 4 + q]);
    iters = std::max(1, int(o->iters));
    {  // whitewater, air field (sim.js:203-206, 221-230)
      const bool on = o->spr_on != 0 && sprayCap > 0, haveSp = o->spr_on != 0;
      p.sprRate = on ? R(o->spr_rate) : R(0);
      p.sprMin = on ? R(std::max(0.05f, o->spr_min)) : R(1);
      p.sprLife = on ? R(o->spr_life) : R(1);
      p.mistFrac = on ? R(o->spr_mist) : R(0);
      p.sprDrag = on ? R(o->spr_drag) : R(1);
      p.mistDrag = on ? R(o->spr_mdrag) : R(4);
      p.vs = o->vs != 0.0f ? R(o->vs) : R(1);
      p.smax = sprayCap;
      p.wwFoamLife = haveSp && o->ww_foam_life != 0.0f ? R(o->ww_foam_life) : R(3);
      p.wwRise = haveSp && o->ww_rise != 0.0f ? R(o->ww_rise) : R(0.3);
      p.wwCrest = haveSp ? R(o->ww_crest) : R(1);
      p.wwImpact = haveSp ? R(o->ww_impact) : R(0.5);
      const bool airo = on && o->air_on != 0;
      if (airo && !airN) {  // airAlloc (sim.js:343-351)
        const double r = std::max(2.0, std::ceil(std::cbrt(double(ncell) / 260000.0)));
        aNx = uint32_t(std::ceil(double(nx) / r)) + 1u;
        aNy = uint32_t(std::ceil(double(ny) / r)) + 1u;
        aNz = uint32_t(std::ceil(double(nz) / r)) + 1u;
        airN = aNx * aNy * aNz;
        airH = hd_ * r;
        air.assign(size_t(airN) * 4, R(0));
        air2.assign(size_t(airN) * 4, R(0));
        airAcc.assign(size_t(airN) * 4, 0);
      }
      p.airOn = airo && airN ? 1u : 0u;
      p.airCouple = airo ? R(o->air_couple) : R(0);
      p.airDecay = airo ? R(o->air_decay) : R(0);
      p.airH = airN ? R(airH) : R(1);
      p.aNx = airN ? aNx : 1u;
      p.aNy = airN ? aNy : 1u;
      p.aNz = airN ? aNz : 1u;
      p.airBlur = airo ? R(o->air_blur) : R(0);
      airOnStep = p.airOn != 0u;
      sprOnStep = on;
      const bool dr = on && o->drip_on != 0;
      p.ndrip = dr ? ndrip : 0u;
      p.dripRate = dr ? R(o->drip_rate) : R(0);
      p.dripDry = dr ? R(o->drip_dry) : R(1);
      dripOnStep = dr && ndrip > 0;
    }
    return ST_OK;
  }
  int runStage(int st) {
    switch (st) {
      case ST_STAGE_SORT: break;
      case ST_STAGE_BLOCKS: stageBlocks(); break;
      case ST_STAGE_CLEAR: if (!onlyTiles) stageClear(); break;
      case ST_STAGE_P2G: if (!onlyTiles) stageP2G(); break;
      case ST_STAGE_NARROWBAND: if (!onlyTiles && nbOn) stageNarrowBand(); break;
      case ST_STAGE_CLASSIFY: if (!onlyTiles) stageClassify(); break;
      case ST_STAGE_NORMFACES: if (!onlyTiles) normFaces(); break;
      case ST_STAGE_DIVERGENCE: if (!onlyTiles) divergence(); break;
      case ST_STAGE_PRESSURE: if (!onlyTiles) stagePressure(); break;
      case ST_STAGE_PROJECT: if (!onlyTiles) { project(); extrapolate(); if (nbOn) saveVel(); } break;
      case ST_STAGE_G2P: if (!onlyTiles) stageG2P(); break;
      case ST_STAGE_SPRAY: if (!onlyTiles && sprayCap) stageSpray(); break;
      case ST_STAGE_RESEED: if (!onlyTiles && nbOn) stageReseed(); break;
      default: return ST_ERR_ARG;
    }
    return ST_OK;
  }
  int step(const st_params* o) {
    int rc = beginStep(o);
    if (rc) return rc;
    for (int s = 0; s < ST_STAGE_COUNT; s++) runStage(s);
    return ST_OK;
  }
  // canonical dumps: tiles with storage ascending, 512 cells each
  std::vector<uint32_t> tilesWithStorage() const {
    std::vector<uint32_t> t;
    for (uint32_t q = 0; q < ntiles; q++) if (tmap[q] != NONE) t.push_back(q);
    return t;
  }
  size_t dumpSize(int what) const {
    size_t nt = tilesWithStorage().size(), C = 512;
    size_t fr = sizeof(R);
    switch (what) {
      case ST_DUMP_TILES: return nt * 4;
      case ST_DUMP_CNT: case ST_DUMP_CTYPE: return nt * C * 4;
      case ST_DUMP_PRES: case ST_DUMP_RHS: return nt * C * fr;
      case ST_DUMP_ACCV: case ST_DUMP_ACCW: return nt * C * 3 * 4;
      case ST_DUMP_FACEV: case ST_DUMP_FACEO: case ST_DUMP_VELG: return nt * C * 3 * fr;
      case ST_DUMP_POS: case ST_DUMP_VEL: return size_t(N) * 4 * fr;
      case ST_DUMP_TMAP: return size_t(ntiles) * 4;
    }
    return 0;
  }
  // test hook: the inverse of dump for the per-cell fields (same canonical order)
  int poke(int what, const void* src, size_t bytes) {
    if (bytes < dumpSize(what)) return ST_ERR_ARG;
    std::vector<uint32_t> ts = tilesWithStorage();
    const char* in = (const char*)src;
    auto perCell = [&](void* dst, size_t esz, size_t per) {
      for (size_t q = 0; q < ts.size(); q++) {
        uint32_t s = tmap[ts[q]];
        std::memcpy((char*)dst + size_t(s) * 512 * per * esz, in + q * 512 * per * esz, 512 * per * esz);
      }
    };
    switch (what) {
      case ST_DUMP_CNT: perCell(cnt.data(), 4, 1); break;
      case ST_DUMP_CTYPE: perCell(ctype.data(), 4, 1); break;
      case ST_DUMP_PRES: perCell(pres.data(), sizeof(R), 1); break;
      case ST_DUMP_RHS: perCell(rhs.data(), sizeof(R), 1); break;
      case ST_DUMP_ACCV: perCell(accVi.data(), 4, 3); break;
      case ST_DUMP_ACCW: perCell(accWi.data(), 4, 3); break;
      case ST_DUMP_FACEV: if constexpr (kF32) perCell(accVi.data(), 4, 3); else perCell(faceV.data(), sizeof(R), 3); break;
      case ST_DUMP_FACEO: if constexpr (kF32) perCell(accWi.data(), 4, 3); else perCell(faceO.data(), sizeof(R), 3); break;
      case ST_DUMP_VELG: perCell(velG.data(), sizeof(R), 3); break;
      default: return ST_ERR_ARG;
    }
    return ST_OK;
  }
  // test hook: one kernel of the narrow band on the current state
  int kernel(int which) {
    switch (which) {
      case 0: phiInitAll(); break;
      case 1: nbPhi(); break;
      case 2: nbJacA(); break;
      case 3: nbJacB(); break;
      case 4: nbVel(); break;
      case 5: saveVel(); break;
      case 6: nbKill(); break;
      case 7: stageReseed(); break;
      case 8: stageNarrowBand(); break;
      case 9: extrapolate(); break;
      default: return ST_ERR_ARG;
    }
    return ST_OK;
  }
  int dump(int what, void* dst, size_t bytes) const {
    if (bytes < dumpSize(what)) return ST_ERR_ARG;
    std::vector<uint32_t> ts = tilesWithStorage();
    char* o = (char*)dst;
    auto perCell = [&](const void* src, size_t esz, size_t per) {
      for (size_t q = 0; q < ts.size(); q++) {
        uint32_t s = tmap[ts[q]];
        std::memcpy(o + q * 512 * per * esz, (const char*)src + size_t(s) * 512 * per * esz, 512 * per * esz);
      }
    };
    switch (what) {
      case ST_DUMP_TILES: std::memcpy(o, ts.data(), ts.size() * 4); break;
      case ST_DUMP_CNT: perCell(cnt.data(), 4, 1); break;
      case ST_DUMP_CTYPE: perCell(ctype.data(), 4, 1); break;
      case ST_DUMP_PRES: perCell(pres.data(), sizeof(R), 1); break;
      case ST_DUMP_RHS: perCell(rhs.data(), sizeof(R), 1); break;
      case ST_DUMP_ACCV: perCell(accVi.data(), 4, 3); break;
      case ST_DUMP_ACCW: perCell(accWi.data(), 4, 3); break;
      case ST_DUMP_FACEV: if constexpr (kF32) perCell(accVi.data(), 4, 3); else perCell(faceV.data(), sizeof(R), 3); break;
      case ST_DUMP_FACEO: if constexpr (kF32) perCell(accWi.data(), 4, 3); else perCell(faceO.data(), sizeof(R), 3); break;
      case ST_DUMP_VELG: perCell(velG.data(), sizeof(R), 3); break;
      case ST_DUMP_POS: std::memcpy(o, pos.data(), size_t(N) * 4 * sizeof(R)); break;
      case ST_DUMP_VEL: std::memcpy(o, vel.data(), size_t(N) * 4 * sizeof(R)); break;
      case ST_DUMP_TMAP: std::memcpy(o, tmap.data(), size_t(ntiles) * 4); break;
      default: return ST_ERR_ARG;
    }
    return ST_OK;
  }
};

struct Handle {
  int fp64;
  Sim<float>* f;
  Sim<double>* d;
};

thread_local std::string g_err;

// the first max(1, min(cap, handed out)) slots, as f64 (4 per point)
template <class R> static uint32_t readSpray(const Sim<R>* s, double* p4, double* v4, uint32_t room) {
  if (!s->sprayCap) return 0;
  uint32_t k = std::max(1u, std::min(s->sprayCap, s->sctl[0]));
  if (k > room) return ~0u;
  for (size_t q = 0; q < size_t(k) * 4; q++) { p4[q] = double(s->sPos[q]); v4[q] = double(s->sVel[q]); }
  return k;
}
}  // namespace

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

int orc_create(const st_opts* o, int fp64, void** out) {
  if (!o || !out || o->d <= 0 || o->max_n == 0) { g_err = "orc_create: bad options"; return ST_ERR_ARG; }
  Handle* h = new Handle{fp64, nullptr, nullptr};
  if (fp64) { h->d = new Sim<double>(); h->d->init(o); }
  else { h->f = new Sim<float>(); h->f->init(o); }
  *out = h;
  return ST_OK;
}
int orc_destroy(void* hp) {
  Handle* h = (Handle*)hp;
  if (!h) return ST_ERR_ARG;
  delete h->f; delete h->d; delete h;
  return ST_OK;
}
#define ORC_DISPATCH(h, expr) ((h)->fp64 ? ((h)->d->expr) : ((h)->f->expr))

int orc_get_dims(void* hp, st_dims* o) {
  Handle* h = (Handle*)hp;
  auto fill = [&](auto* s) {
    o->d = float(s->d); o->h = float(s->h); o->ratio = float(s->ratio); o->rest = float(s->rest);
    o->x0 = float(s->bx0); o->y0 = float(s->by0); o->z0 = float(s->bz0); o->x1 = float(s->bx1); o->y1 = float(s->by1); o->z1 = float(s->bz1);
    o->nx = s->nx; o->ny = s->ny; o->nz = s->nz; o->tx = int32_t(s->tx); o->ty = int32_t(s->ty); o->tz = int32_t(s->tz);
    o->ntiles = s->ntiles; o->scap = s->scap; o->scap_max = s->scapMax; o->max_n = s->maxn; o->ncell = s->ncell;
  };
  if (h->fp64) fill(h->d); else fill(h->f);
  return ST_OK;
}
int orc_upload_particles(void* hp, uint32_t first, uint32_t count, const float* p4, const float* v4) {
  Handle* h = (Handle*)hp;
  auto up = [&](auto* s) -> int {
    if (size_t(first) + count > s->maxn) return ST_ERR_ARG;
    for (size_t q = 0; q < size_t(count) * 4; q++) {
      s->pos[size_t(first) * 4 + q] = p4[q];
      s->vel[size_t(first) * 4 + q] = v4[q];
    }
    s->N = std::max(s->N, first + count);
    return ST_OK;
  };
  return h->fp64 ? up(h->d) : up(h->f);
}
int orc_set_count(void* hp, uint32_t n) {
  Handle* h = (Handle*)hp;
  if (h->fp64) h->d->N = std::min(n, h->d->maxn); else h->f->N = std::min(n, h->f->maxn);
  return ST_OK;
}
int orc_set_colliders(void* hp, const float* cf, uint32_t nc, const float* sdf, size_t nsdf, uint64_t sig) {
  Handle* h = (Handle*)hp;
  auto set = [&](auto* s) {
    using RR = typename std::remove_reference<decltype(s->cols[0])>::type;
    s->cols.assign(size_t(nc) * 32, RR(0));
    for (size_t q = 0; q < size_t(nc) * 32; q++) s->cols[q] = RR(cf[q]);
    if (sig != s->sdfSig || s->sdf.size() != nsdf) {
      s->sdf.assign(nsdf, RR(0));
      for (size_t q = 0; q < nsdf; q++) s->sdf[q] = RR(sdf[q]);
      s->sdfSig = sig;
    }
    s->ncolsTotal = nc;
  };
  if (h->fp64) set(h->d); else set(h->f);
  return ST_OK;
}
int orc_set_drips(void* hp, const float* a, uint32_t n) {
  Handle* h = (Handle*)hp;
  if (h->fp64) h->d->setDrips(a, n); else h->f->setDrips(a, n);
  return ST_OK;
}
int orc_read_deep(void* hp, double* out, uint32_t room, uint32_t* n) {
  Handle* h = (Handle*)hp;
  *n = h->fp64 ? h->d->readDeep(out, room) : h->f->readDeep(out, room);
  return ST_OK;
}
int orc_set_spray_cap(void* hp, uint32_t n) {
  Handle* h = (Handle*)hp;
  if (h->fp64) h->d->setSprayCap(n); else h->f->setSprayCap(n);
  return ST_OK;
}
// test hooks: put whitewater points into slots 0..n-1 and set the ring counter;
// read the air field (4 values per air cell)
int orc_spray_put(void* hp, const double* p4, const double* v4, uint32_t n, uint32_t handed) {
  Handle* h = (Handle*)hp;
  auto put = [&](auto* s) {
    if (n > s->sprayCap) return ST_ERR_ARG;
    for (size_t q = 0; q < size_t(n) * 4; q++) { s->sPos[q] = p4[q]; s->sVel[q] = v4[q]; }
    s->sctl[0] = handed;
    return ST_OK;
  };
  int rc = h->fp64 ? put(h->d) : put(h->f);
  if (rc) g_err = "orc_spray_put: more points than the pool holds";
  return rc;
}
int orc_read_air(void* hp, double* out, uint32_t room, uint32_t* n, double* cell) {
  Handle* h = (Handle*)hp;
  auto rd = [&](auto* s) {
    *n = s->airN;
    *cell = s->airH;
    if (s->airN > room) return ST_ERR_ARG;
    for (size_t q = 0; q < size_t(s->airN) * 4; q++) out[q] = double(s->air[q]);
    return ST_OK;
  };
  return h->fp64 ? rd(h->d) : rd(h->f);
}
void orc_sincosf(float x, float* sn, float* cs) { st_sincosf(x, sn, cs); }
int orc_read_spray(void* hp, double* p4, double* v4, uint32_t room, uint32_t* n) {
  Handle* h = (Handle*)hp;
  uint32_t k = h->fp64 ? readSpray(h->d, p4, v4, room) : readSpray(h->f, p4, v4, room);
  if (k == ~0u) { g_err = "orc_read_spray: output arrays too small"; return ST_ERR_ARG; }
  *n = k;
  return ST_OK;
}
int orc_begin_step(void* hp, const st_params* p) { Handle* h = (Handle*)hp; return ORC_DISPATCH(h, beginStep(p)); }
int orc_stage(void* hp, int32_t st) { Handle* h = (Handle*)hp; return ORC_DISPATCH(h, runStage(st)); }
int orc_step(void* hp, const st_params* p) { Handle* h = (Handle*)hp; return ORC_DISPATCH(h, step(p)); }
int orc_grow(void* hp, uint32_t n) { Handle* h = (Handle*)hp; return ORC_DISPATCH(h, grow(n)); }
int orc_read_ctl(void* hp, uint32_t ctl[16]) {
  Handle* h = (Handle*)hp;
  std::memcpy(ctl, h->fp64 ? h->d->ctl : h->f->ctl, 64);
  return ST_OK;
}
int orc_dump_size(void* hp, int32_t what, size_t* bytes) { Handle* h = (Handle*)hp; *bytes = ORC_DISPATCH(h, dumpSize(what)); return ST_OK; }
int orc_dump(void* hp, int32_t what, void* dst, size_t bytes) { Handle* h = (Handle*)hp; return ORC_DISPATCH(h, dump(what, dst, bytes)); }
int orc_poke(void* hp, int32_t what, const void* src, size_t bytes) { Handle* h = (Handle*)hp; return ORC_DISPATCH(h, poke(what, src, bytes)); }
int orc_kernel(void* hp, int32_t which) { Handle* h = (Handle*)hp; return ORC_DISPATCH(h, kernel(which)); }
float orc_expf(float x) { return st_expf(x); }
// test hooks: the gust noise of g2p (vnz) and its lattice hash (hsh3), in f32 or f64
double orc_vnz(double x, double y, double z, int fp64) {
  if (fp64) return Sim<double>::vnz(v3(x, y, z));
  return double(Sim<float>::vnz(v3(float(x), float(y), float(z))));
}
double orc_hsh3(double x, double y, double z, int fp64) {
  if (fp64) return Sim<double>::hsh3(v3(x, y, z));
  return double(Sim<float>::hsh3(v3(float(x), float(y), float(z))));
}
// test hook: the ocean helpers of the begun step at n points (x, y, z, -) ->
// (oceanEta(x, z), oceanVel(p) x, y, z, layerW(p)) per point
int orc_ocean_eval(void* hp, const double* pts4, uint32_t n, double* out5) {
  Handle* h = (Handle*)hp;
  auto ev = [&](auto* s) {
    using RR = typename std::remove_reference<decltype(s->h)>::type;
    for (size_t q = 0; q < n; q++) {
      V3<RR> p = v3(RR(pts4[q * 4]), RR(pts4[q * 4 + 1]), RR(pts4[q * 4 + 2]));
      V3<RR> v = s->oceanVel(p);
      out5[q * 5] = double(s->oceanEta(p.x, p.z));
      out5[q * 5 + 1] = double(v.x);
      out5[q * 5 + 2] = double(v.y);
      out5[q * 5 + 3] = double(v.z);
      out5[q * 5 + 4] = double(s->layerW(p));
    }
  };
  if (h->fp64) ev(h->d); else ev(h->f);
  return ST_OK;
}
int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
}

// flip_device.h - device-side parameter block and helpers of the FLIP substep.
//
// Sparse grid: the box is cut into tiles of 8x8x8 cells, t = ti + tx*(tj + ty*tk);
// tiles covering the box plus one cell on each high side (so the faces on the far
// walls have an owner cell).  A tile near particles holds a slot of the block pool;
// tmap[t] is that slot or NONE.  Per slot, 512 cells in local order
// i&7 + 8*((j&7) + 8*(k&7)); each cell owns its lower x, y, z faces.  Slot 0 is a
// spare that every tile without storage reads (cleared each step, never
// allocated), as in the reference (src/flip/wgsl-core.js:79-89).
//
// All arithmetic is f32 with contraction off: every rounding is the one the
// expressions below write, in the order the reference's WGSL writes them.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "st_math.h"

#define ST_NONE 0xffffffffu
#define ST_DEADY (-1.0e5f)
#define ST_MAX_COLS 32
// p2g's window around a tile is 10^3 cells; the 488 outside the tile (the halo
// shell) carry 3 faces x (velocity sum, weight sum) each
#define P2G_HALO_CELLS 488
#define P2G_HALO_WORDS (P2G_HALO_CELLS * 6)

// The uniform block of the reference (src/flip/wgsl-core.js:3-20, packed by
// FlipSim.step at src/flip/sim.js:196-231), minus the display, spray, air and
// ocean words that the substep does not read.
struct DParams {
  uint32_t n, nx, ny, nz, ncols, open, ncell, scap;
  float dt, h, ih, gravity, flip, rest, comp, omega;
  float x0, y0, z0, vmax, margin, scale, x1, y1;
  float z1, q0, q1, q2, q3, q4, q5, q6;
  float q7, q8, q9, q10, q11, q12, sheetMin, band;
  uint32_t tx, ty, tz, lcap, seed, nb, nbInit, nfill;
  uint32_t nfv, ntiles;
  float settle;  // smoothstep(0.1, 0.5, tm), the same for every particle of the step
  uint32_t sorTag;  // resident SOR: tag of the substep's first halo exchange (the k-th carries sorTag + k)
  uint32_t iters;   // SOR iterations of the substep
  float fills[16][4];
  // whitewater and the air field (wgsl-core.js:10-16; sim.js:203-206, 221-230)
  float sprRate, sprMin, sprLife, mistFrac, sprDrag, mistDrag, vs;
  uint32_t smax;      // whitewater slots (0: no spray stage)
  float wwFoamLife, wwRise, wwCrest, wwImpact;
  uint32_t airOn;
  float airCouple, airDecay, airH;
  uint32_t aNx, aNy, aNz;
  float airBlur;
  uint32_t sprOn;     // new points are thrown off this step
  uint32_t ndrip;     // drip samples on the colliders (0: no drips this step)
  float dripRate, dripDry;
  // the ocean side boundary (wgsl-core.js:12-13, 201-226; sim.js:207-218)
  uint32_t ocean, nwave;
  uint32_t rk2, extrap;  // opt-in modes (0: the reference's path)
  float otime, olevel, olayer, odepth, oy0;
  float oease;        // 1 - exp(-8 dt): g2p's easing toward the waves per step
  // squared lengths of vn - vo (v0 - vo) at or below which g2p's impact (turb)
  // quotient cannot exceed 15 (0.35), so both are 0 there
  float impS2, turbS2;
  float waves[64][4]; // per wave two rows: (kx, kz, amplitude, angular frequency), (phase, wavenumber, 0, 0)
};

// Collider records: 8 float4 per collider (gpu.js:3-4): (sdf offset, nx, ny, nz),
// (origin, cell), the three rows of R^T, position, velocity, angular velocity.
struct DCols {
  float4 c[ST_MAX_COLS * 8];
  // per collider, from its record: (lo, hi margin) of colNear, lo = -1 - q0 / cell,
  // hq = q0 / cell (the same f32 divisions, done once on the host)
  float4 near[ST_MAX_COLS];
  // per collider, a world box holding the whole colNear region with room to
  // spare (the host maps the local near box through the inverse rotation, in
  // f64, and pads it), or +-inf when it cannot be made: a point outside it
  // is not near, which colNear then need not work out
  float4 wlo[ST_MAX_COLS];
  float4 whi[ST_MAX_COLS];
};

// Device pointers of one sim, passed by value to every kernel.
struct DGrid {
  float4* pos;
  float4* vel;
  uint32_t* pid;
  float4* posB;  // the other particle buffer: g2p writes the next state here in bin order, then
                 // the host swaps it with pos/vel/pid
  float4* velB;
  uint32_t* pidB;
  uint32_t* binKey;    // per particle (A order): bin = tile of its cell, ntiles if dead
  uint32_t* binIdx;    // bin order: index into pos/vel/pid of the k-th particle
  uint32_t* binRank;   // per particle (A order): its place within its bin
  uint32_t* binCount;  // per tile (+1 dead bin): particles whose cell is in the tile
  uint32_t* binStart;  // exclusive prefix of binCount (ntiles + 2 entries, last = N)
  uint32_t* scanTmp;   // per-1024-block sums of the bin scan
  uint32_t* tmap;
  uint32_t* need;
  uint32_t* newList;   // pairs: tile, slot (or NONE)
  uint32_t* slotTile;
  uint32_t* fstack;
  uint32_t* slotList;
  uint32_t* nbr;       // per slot: slots of the 27 tiles around it (0 none, WSLOT_OUT outside)
  uint32_t* nbrL;      // the same rows in slotList order (row b for slotList[b])
  uint32_t* wg;        // per-workgroup scan scratch
  uint32_t* part;      // per-workgroup statistics partials (vmax; fluid cells per colour)
  uint32_t* ctl;       // control words, as in the reference's ctl
  int32_t* cnt;
  uint32_t* ctype;
  float* pres;
  float* pres2;        // SOR ping-pong partner of pres
  float* rhs;
  int32_t* accV;       // 3 per cell: fixed-point sums, then new face velocity bits
  int32_t* accW;       // 3 per cell: fixed-point weights, then old face velocity bits
  int32_t* p2gHalo;    // per slot, P2G_HALO_WORDS: the sums p2g made for the halo shell of the
                       // slot's window (cells of neighbour tiles), which the owners add in
  float* velG;         // 3 per cell (narrow band)
  uint8_t* sten;       // per cell: SOR stencil byte (bit 7 fluid, bits 0-2 non-solid neighbours)
  const float* sdf;
  const DCols* cols;
  float4* scratch4;    // readback scratch (id order)
  float4* foamTab;     // per step, for a density sum c < FOAM_TAB: the g2p factors of c (d_foamRow)
  // resident SOR
  uint32_t* mortonTiles; // tile indices in Morton (z-order) order of their tile coordinates
  uint32_t* coopFlags;   // per Morton position: 1 when the tile is in use
  uint32_t* coopOffs;    // exclusive prefix of coopFlags
  uint32_t* coopList;    // slots in use, in Morton order (per slot)
  uint64_t* sorHalo;     // per colour, slot and face: the 32 face cells of that colour, each with the tag of
                         // the half-sweep that wrote it (2 x slots x 6 x 32 pairs: value bits low, tag high)
  uint32_t* sorBar;      // per resident-SOR workgroup: the last barrier epoch it reached
  // whitewater (setSprayCap, sim.js:333-340): smax slots of points, control words, and the air field
  float4* sPos;          // xyz, seconds of life left (0: free)
  float4* sVel;          // velocity, kind (0 spray, 1 mist, 2 foam, 3 bubble) + 0.4 x how long it has had it
  float4* sAux;          // water velocity at the point (xyz), how wet it is (w)
  uint32_t* sctl;        // [0] slots handed out (a ring counter)
  uint32_t* sprCnt;      // per particle id: points it throws off this step; then the exclusive prefix (N + 1)
  uint32_t* sprOff;
  uint32_t* sprSums;     // per-1024 block sums of that scan
  float4* air;           // air field cell velocities, and the next step's
  float4* air2;
  int32_t* airAcc;       // per air cell: fixed-point (x 256) velocity and weight pushed into it
  // drips (dripUpdate, sim.js:353-363): per sample two float4 -- local position and collider number,
  // local normal and wetness -- and the flags / scan that place this step's drops
  float4* drip;
  uint32_t* dripFlag;
  uint32_t* dripOff;
  uint32_t* dripSums;
};

// ---------------------------------------------------------------------------
// sparse addressing (wgsl-core.js:76-112)

__device__ __forceinline__ uint32_t d_tileOf(const DParams& P, int32_t i, int32_t j, int32_t k) {
  // every product stays below the tile count, which st_create keeps below 2^24
  return uint32_t(i >> 3) + __umul24(P.tx, uint32_t(j >> 3) + __umul24(P.ty, uint32_t(k >> 3)));
}
__device__ __forceinline__ uint32_t d_cellIdx(const DParams& P, const uint32_t* __restrict__ tmap, int32_t i, int32_t j, int32_t k) {
  uint32_t s = tmap[d_tileOf(P, i, j, k)];
  return (s == ST_NONE ? 0u : s) * 512u + uint32_t(i & 7) + 8u * (uint32_t(j & 7) + 8u * uint32_t(k & 7));
}
__device__ __forceinline__ int32_t d_iclamp(int32_t x, int32_t lo, int32_t hi) { return min(max(x, lo), hi); }
__device__ __forceinline__ uint32_t d_faceIndex(const DParams& P, const uint32_t* __restrict__ tmap, uint32_t axis, int32_t i, int32_t j, int32_t k) {
  int32_t X = int32_t(P.nx), Y = int32_t(P.ny), Z = int32_t(P.nz);
  if (axis == 0) return d_cellIdx(P, tmap, d_iclamp(i, 0, X), d_iclamp(j, 0, Y - 1), d_iclamp(k, 0, Z - 1)) * 3u;
  if (axis == 1) return d_cellIdx(P, tmap, d_iclamp(i, 0, X - 1), d_iclamp(j, 0, Y), d_iclamp(k, 0, Z - 1)) * 3u + 1u;
  return d_cellIdx(P, tmap, d_iclamp(i, 0, X - 1), d_iclamp(j, 0, Y - 1), d_iclamp(k, 0, Z)) * 3u + 2u;
}
__device__ __forceinline__ int3 d_slotCoord(const DParams& P, const uint32_t* __restrict__ slotTile, uint32_t s, uint32_t local) {
  uint32_t t = slotTile[s], tt = t / P.tx;
  return make_int3(int32_t((t % P.tx) * 8u + (local & 7u)), int32_t((tt % P.ty) * 8u + ((local >> 3) & 7u)),
                   int32_t((tt / P.ty) * 8u + (local >> 6)));
}
__device__ __forceinline__ uint32_t d_wallType(const DParams& P, int32_t i, int32_t j, int32_t k, bool& outside) {
  outside = true;
  if (i < 0) return (P.open & 1u) ? 0u : 2u;
  if (i >= int32_t(P.nx)) return (P.open & 2u) ? 0u : 2u;
  if (j < 0) return (P.open & 4u) ? 0u : 2u;
  if (j >= int32_t(P.ny)) return (P.open & 8u) ? 0u : 2u;
  if (k < 0) return (P.open & 16u) ? 0u : 2u;
  if (k >= int32_t(P.nz)) return (P.open & 32u) ? 0u : 2u;
  outside = false;
  return 0u;
}
__device__ __forceinline__ uint32_t d_ctypeOf(const DParams& P, const uint32_t* __restrict__ tmap, const uint32_t* __restrict__ ctype,
                                              int32_t i, int32_t j, int32_t k) {
  bool out;
  uint32_t w = d_wallType(P, i, j, k, out);
  if (out) return w;
  return ctype[d_cellIdx(P, tmap, i, j, k)] & 0xffu;
}
__device__ __forceinline__ bool d_inDom(const DParams& P, int3 p) {
  return p.x >= 0 && p.y >= 0 && p.z >= 0 && p.x < int32_t(P.nx) && p.y < int32_t(P.ny) && p.z < int32_t(P.nz);
}
__device__ __forceinline__ float3 d_cellCenter(const DParams& P, int32_t i, int32_t j, int32_t k) {
  return make_float3(P.x0 + (float(i) + 0.5f) * P.h, P.y0 + (float(j) + 0.5f) * P.h, P.z0 + (float(k) + 0.5f) * P.h);
}
__device__ __forceinline__ float3 d_facePos(const DParams& P, int3 p, uint32_t axis) {
  return make_float3(P.x0 + (float(p.x) + (axis == 0 ? 0.0f : 0.5f)) * P.h, P.y0 + (float(p.y) + (axis == 1 ? 0.0f : 0.5f)) * P.h,
                     P.z0 + (float(p.z) + (axis == 2 ? 0.0f : 0.5f)) * P.h);
}
__device__ __forceinline__ int3 d_particleCell(const DParams& P, float3 p) {
  int32_t cx = st_f2i_sat(floorf((p.x - P.x0) * P.ih));
  int32_t cy = st_f2i_sat(floorf((p.y - P.y0) * P.ih));
  int32_t cz = st_f2i_sat(floorf((p.z - P.z0) * P.ih));
  return make_int3(d_iclamp(cx, 0, int32_t(P.nx) - 1), d_iclamp(cy, 0, int32_t(P.ny) - 1), d_iclamp(cz, 0, int32_t(P.nz) - 1));
}

// ---------------------------------------------------------------------------
// small f32 vector helpers, evaluated left to right as the WGSL writes them

__device__ __forceinline__ float3 f3sub(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 f3add(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 f3mul(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float3 f3div(float3 a, float s) { return make_float3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ float f3dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float f3len(float3 a) { return sqrtf(f3dot(a, a)); }
__device__ __forceinline__ float f3get(float3 a, uint32_t axis) { return axis == 0 ? a.x : (axis == 1 ? a.y : a.z); }
__device__ __forceinline__ float3 c3(float4 v) { return make_float3(v.x, v.y, v.z); }

// ---------------------------------------------------------------------------
// colliders (gpu.js:6-38, wgsl-core.js:114-131)

__device__ __forceinline__ void d_sdfEval(const DCols* __restrict__ C, const float* __restrict__ sdf, uint32_t k, float3 p, float& phiOut, float3& wnOut) {
  float4 a = C->c[k * 8u], b = C->c[k * 8u + 1u];
  float3 r0 = c3(C->c[k * 8u + 2u]), r1 = c3(C->c[k * 8u + 3u]), r2 = c3(C->c[k * 8u + 4u]);
  float3 d = f3sub(p, c3(C->c[k * 8u + 5u]));
  float3 lp = make_float3(f3dot(r0, d), f3dot(r1, d), f3dot(r2, d));
  float3 g = f3div(f3sub(lp, make_float3(b.x, b.y, b.z)), b.w);
  const float e = 1.0001f;
  float3 gc = make_float3(st_clampf(g.x, 0.0f, a.y - e), st_clampf(g.y, 0.0f, a.z - e), st_clampf(g.z, 0.0f, a.w - e));
  float3 fl = make_float3(floorf(gc.x), floorf(gc.y), floorf(gc.z));
  float3 f = f3sub(gc, fl);
  uint32_t snx = st_f2u_sat(a.y), sny = st_f2u_sat(a.z);
  uint32_t base = st_f2u_sat(a.x) + st_f2u_sat(fl.x) + snx * (st_f2u_sat(fl.y) + sny * st_f2u_sat(fl.z));
  uint32_t s = snx * sny;
  float v000 = sdf[base], v100 = sdf[base + 1u], v010 = sdf[base + snx], v110 = sdf[base + snx + 1u];
  float v001 = sdf[base + s], v101 = sdf[base + s + 1u], v011 = sdf[base + s + snx], v111 = sdf[base + s + snx + 1u];
  float c00 = st_mixf(v000, v100, f.x), c10 = st_mixf(v010, v110, f.x), c01 = st_mixf(v001, v101, f.x), c11 = st_mixf(v011, v111, f.x);
  float c0 = st_mixf(c00, c10, f.y), c1 = st_mixf(c01, c11, f.y);
  float phi = st_mixf(c0, c1, f.z);
  float3 out = f3sub(g, gc);
  float3 gl;
  if (f3dot(out, out) > 0.0f) {
    float l = f3len(out);
    phi = phi + l * b.w;
    gl = f3div(out, l);
  } else {
    gl = make_float3(st_mixf(st_mixf(v100 - v000, v110 - v010, f.y), st_mixf(v101 - v001, v111 - v011, f.y), f.z),
                     st_mixf(c10 - c00, c11 - c01, f.z), c1 - c0);
    float l = f3len(gl);
    gl = f3div(gl, fmaxf(l, 1e-9f));
  }
  wnOut = f3add(f3add(f3mul(r0, gl.x), f3mul(r1, gl.y)), f3mul(r2, gl.z));
  phiOut = phi;
}
__device__ __forceinline__ float3 d_colSurfVel(const DCols* __restrict__ C, uint32_t k, float3 p) {
  float3 c = c3(C->c[k * 8u + 5u]), v = c3(C->c[k * 8u + 6u]), w = c3(C->c[k * 8u + 7u]);
  float3 q = f3sub(p, c);
  float3 cr = make_float3(w.y * q.z - w.z * q.y, w.z * q.x - w.x * q.z, w.x * q.y - w.y * q.x);
  return f3add(v, cr);
}
__device__ __forceinline__ bool d_colNear(const DParams& P, const DCols* __restrict__ C, uint32_t k, float3 p) {
  if (P.q10 > 0.5f) return true;
  const float4 wl = C->wlo[k], wh = C->whi[k];
  if (p.x < wl.x || p.y < wl.y || p.z < wl.z || p.x > wh.x || p.y > wh.y || p.z > wh.z) return false;
  float4 b = C->c[k * 8u + 1u];
  float3 d = f3sub(p, c3(C->c[k * 8u + 5u]));
  float3 lp = make_float3(f3dot(c3(C->c[k * 8u + 2u]), d), f3dot(c3(C->c[k * 8u + 3u]), d), f3dot(c3(C->c[k * 8u + 4u]), d));
  float3 g = f3div(f3sub(lp, make_float3(b.x, b.y, b.z)), b.w);
  float4 a = C->c[k * 8u];
  const float lo = C->near[k].x, hq = C->near[k].y;
  return g.x > lo && g.y > lo && g.z > lo && g.x < a.y + hq && g.y < a.z + hq && g.z < a.w + hq;
}
__device__ __forceinline__ int32_t d_colAt(const DParams& P, const DCols* __restrict__ C, const float* __restrict__ sdf, float3 p) {
  for (uint32_t k = 0; k < P.ncols; k++) {
    if (!d_colNear(P, C, k, p)) continue;
    float phi;
    float3 wn;
    d_sdfEval(C, sdf, k, p, phi, wn);
    if (phi < P.q0) return int32_t(k);
  }
  return -1;
}
__device__ __forceinline__ int32_t d_colAtCell(const DParams& P, const uint32_t* __restrict__ tmap, const uint32_t* __restrict__ ctype, int32_t i, int32_t j, int32_t k) {
  if (i < 0 || j < 0 || k < 0 || i >= int32_t(P.nx) || j >= int32_t(P.ny) || k >= int32_t(P.nz)) return -1;
  return int32_t((ctype[d_cellIdx(P, tmap, i, j, k)] >> 8) & 0xffu) - 1;
}

// ---------------------------------------------------------------------------
// narrow band helpers (wgsl-core.js:160-200, 228-238)

__device__ __forceinline__ float d_phiOfBits(uint32_t ct) { return 8.0f - float(ct >> 16) / 4096.0f; }
__device__ __forceinline__ uint32_t d_phiPutBits(uint32_t ct, float v) {
  uint32_t q = st_f2u_sat(st_clampf(rintf((8.0f - v) * 4096.0f), 0.0f, 65535.0f));
  return (ct & 0xffffu) | (q << 16);
}
__device__ __forceinline__ bool d_calmV(const DParams& P, float3 v) { return f3len(v) < 9.0f * P.h; }

// ---------------------------------------------------------------------------
// ocean (wgsl-core.js:201-226): a sum of linear waves on water of depth odepth
// over a floor at oy0, imposed through a layer at the side walls

// surface height at (x, z)
__device__ __forceinline__ float d_oceanEta(const DParams& P, float x, float z) {
  float e = 0.0f;
  for (uint32_t k = 0; k < P.nwave; k++) {
    const float* A = P.waves[k * 2u];
    const float* B = P.waves[k * 2u + 1u];
    float sn, cs;
    st_sincosf_wide(A[0] * x + A[1] * z - A[3] * P.otime + B[0], &sn, &cs);
    e = e + A[2] * cs;
  }
  return P.olevel + e;
}
// water velocity under the waves; the height is stretched so the crest gets the
// surface value (Wheeler stretching)
__device__ __forceinline__ float3 d_oceanVel(const DParams& P, float3 p) {
  const float D = fmaxf(P.odepth, 1.0e-3f);
  const float eta = d_oceanEta(P, p.x, p.z);
  const float ys = st_clampf((p.y - P.oy0) * D / fmaxf(D + eta - P.olevel, 0.1f * D), 0.0f, D);
  float3 v = make_float3(0.0f, 0.0f, 0.0f);
  for (uint32_t k = 0; k < P.nwave; k++) {
    const float* A = P.waves[k * 2u];
    const float* B = P.waves[k * 2u + 1u];
    const float kk = fmaxf(B[1], 1.0e-4f);
    float sn, cs;
    st_sincosf_wide(A[0] * p.x + A[1] * p.z - A[3] * P.otime + B[0], &sn, &cs);
    const float e1 = st_expf(kk * (ys - D));
    const float e2 = st_expf(-kk * (ys + D));
    const float dn = fmaxf(1.0f - st_expf(-2.0f * kk * D), 1.0e-4f);
    const float aw = A[2] * A[3];
    const float hc = aw * (e1 + e2) / dn * cs / kk;
    v = f3add(v, make_float3(hc * A[0], aw * (e1 - e2) / dn * sn, hc * A[1]));
  }
  return v;
}
// 1 at the side walls, 0 at the layer's width in from them (0 without the ocean)
__device__ __forceinline__ float d_layerW(const DParams& P, float3 p) {
  if (P.ocean == 0u) return 0.0f;
  const float dw = fminf(fminf(p.x - P.x0, P.x1 - p.x), fminf(p.z - P.z0, P.z1 - p.z));
  return st_clampf(1.0f - dw / fmaxf(P.olayer, 1.0e-4f), 0.0f, 1.0f);
}

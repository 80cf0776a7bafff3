// flip_kernels.hip - HIP kernels of the FLIP substep for gfx950 (wave64).
//
// One kernel per step of the reference's stage list (src/flip/sim.js:245-310),
// except where noted.  Work over active cells is a grid-stride loop over the
// device-side count of tiles in use (ctl[4]), so no launch needs a host readback
// and every launch can be captured in a graph.
#include "flip_device.h"
#include "flip_kernels.h"

#define WG 256

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

// ---------------------------------------------------------------------------
// generic fills (no hipMemsetAsync, so the launches can live in a graph)

__global__ void k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------------------
// blocks (wgsl-blocks.js:4-60)

// tileMark + markAround (wgsl-blocks.js:10-14, wgsl-core.js:239-246): every tile
// within two cells of a live particle is needed.  All writers store 1.
__global__ void k_tile_mark(const DParams* __restrict__ PP, DGrid G, const float4* __restrict__ pos) {
  const DParams& P = *PP;
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  float4 q = pos[i];
  if (q.y < ST_DEADY) return;
  int3 cc = d_particleCell(P, make_float3(q.x, q.y, q.z));
  int32_t lx = max(cc.x - 2, 0) >> 3, ly = max(cc.y - 2, 0) >> 3, lz = max(cc.z - 2, 0) >> 3;
  int32_t hx = min(cc.x + 2, int32_t(P.tx) * 8 - 1) >> 3, hy = min(cc.y + 2, int32_t(P.ty) * 8 - 1) >> 3, hz = min(cc.z + 2, int32_t(P.tz) * 8 - 1) >> 3;
  for (int32_t tk = lz; tk <= hz; tk++)
    for (int32_t tj = ly; tj <= hy; tj++)
      for (int32_t ti = lx; ti <= hx; ti++) {
        uint32_t t = uint32_t(ti) + P.tx * (uint32_t(tj) + P.ty * uint32_t(tk));
        if (G.need[t] == 0u) G.need[t] = 1u;
      }
}

// Flags of one tile: freed (no longer needed, had a slot), new (needed, had
// none), kept (needed, had one).  tileFree in wgsl-blocks.js:16-21.
struct TileFlags {
  bool fr, nw, kp;
};
__device__ __forceinline__ TileFlags tile_flags(const DGrid& G, uint32_t t, uint32_t ntiles) {
  TileFlags f{false, false, false};
  if (t < ntiles) {
    uint32_t nd = G.need[t], s = G.tmap[t];
    f.fr = nd == 0u && s != ST_NONE;
    f.nw = nd != 0u && s == ST_NONE;
    f.kp = nd != 0u && s != ST_NONE;
  }
  return f;
}

// workgroup-wide exclusive ranks of three flags (256 threads = 4 waves)
struct Ranks3 {
  uint32_t r[3];
  uint32_t tot[3];
};
__device__ __forceinline__ Ranks3 wg_ranks3(bool a, bool b, bool c) {
  __shared__ uint32_t wsum[3][WG / 64];
  uint64_t m[3] = {__ballot(a), __ballot(b), __ballot(c)};
  uint32_t w = threadIdx.x >> 6;
  if (lane_id() == 0)
    for (int q = 0; q < 3; q++) wsum[q][w] = uint32_t(__popcll(m[q]));
  __syncthreads();
  Ranks3 R;
  for (int q = 0; q < 3; q++) {
    uint32_t before = 0, tot = 0;
    for (uint32_t v = 0; v < WG / 64; v++) {
      if (v < w) before += wsum[q][v];
      tot += wsum[q][v];
    }
    R.r[q] = before + uint32_t(__popcll(m[q] & lanemask_lt()));
    R.tot[q] = tot;
  }
  __syncthreads();
  return R;
}

__global__ void __launch_bounds__(WG) k_tile_count(const DParams* __restrict__ PP, DGrid G) {
  const DParams& P = *PP;
  uint32_t t = blockIdx.x * WG + threadIdx.x;
  TileFlags f = tile_flags(G, t, P.ntiles);
  Ranks3 R = wg_ranks3(f.fr, f.nw, f.kp);
  if (threadIdx.x == 0) {
    G.wg[blockIdx.x * 3 + 0] = R.tot[0];
    G.wg[blockIdx.x * 3 + 1] = R.tot[1];
    G.wg[blockIdx.x * 3 + 2] = R.tot[2];
  }
}

// One workgroup: exclusive scan of the per-workgroup counts, then the stack
// arithmetic of tileFree/tileAlloc/tileFix (wgsl-blocks.js:16-35) in closed form:
// the k-th new tile (tile order) pops stack entry top1-1-k, so the first
// min(new, top1) new tiles get slots and the rest count as no room.
__global__ void __launch_bounds__(1024) k_tile_scan(const DParams* __restrict__ PP, DGrid G, uint32_t nwg) {
  const DParams& P = *PP;
  __shared__ uint32_t part[3][1024 / 64];
  __shared__ uint32_t carry[3];
  if (threadIdx.x < 3) carry[threadIdx.x] = 0;
  __syncthreads();
  for (uint32_t base = 0; base < nwg; base += 1024) {
    uint32_t i = base + threadIdx.x;
    uint32_t v[3];
    for (int q = 0; q < 3; q++) v[q] = i < nwg ? G.wg[i * 3 + q] : 0u;
    // inclusive wave scan
    uint32_t s[3] = {v[0], v[1], v[2]};
    for (int off = 1; off < 64; off <<= 1)
      for (int q = 0; q < 3; q++) {
        uint32_t o = __shfl_up(s[q], off);
        if (int(lane_id()) >= off) s[q] += o;
      }
    uint32_t w = threadIdx.x >> 6;
    if (lane_id() == 63)
      for (int q = 0; q < 3; q++) part[q][w] = s[q];
    __syncthreads();
    uint32_t before[3] = {0, 0, 0}, tot[3] = {0, 0, 0};
    for (uint32_t u = 0; u < 1024 / 64; u++)
      for (int q = 0; q < 3; q++) {
        if (u < w) before[q] += part[q][u];
        tot[q] += part[q][u];
      }
    if (i < nwg)
      for (int q = 0; q < 3; q++) G.wg[i * 3 + q] = carry[q] + before[q] + s[q] - v[q];
    __syncthreads();
    if (threadIdx.x < 3) carry[threadIdx.x] += tot[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    uint32_t F = carry[0], Nn = carry[1], K = carry[2];
    uint32_t top0 = G.ctl[2];
    uint32_t top1 = top0 + F;
    uint32_t nalloc = min(Nn, top1);
    G.ctl[2] = top1 - nalloc;
    G.ctl[3] = Nn;
    G.ctl[4] = K + nalloc;
    G.ctl[5] = Nn - nalloc;
    G.ctl[6] = 0u;
    G.ctl[12] = top0;
    G.ctl[13] = top1;
    G.ctl[14] = nalloc;
  }
  (void)P;
}

// Freed tiles push their slot (tile order), new tiles are listed (tile order),
// kept tiles take their place in slotList (tile order).
__global__ void __launch_bounds__(WG) k_tile_scatter(const DParams* __restrict__ PP, DGrid G) {
  const DParams& P = *PP;
  uint32_t t = blockIdx.x * WG + threadIdx.x;
  TileFlags f = tile_flags(G, t, P.ntiles);
  Ranks3 R = wg_ranks3(f.fr, f.nw, f.kp);
  uint32_t baseF = G.wg[blockIdx.x * 3 + 0], baseN = G.wg[blockIdx.x * 3 + 1], baseK = G.wg[blockIdx.x * 3 + 2];
  uint32_t top0 = G.ctl[12], nalloc = G.ctl[14];
  if (f.fr) {
    G.fstack[top0 + baseF + R.r[0]] = G.tmap[t];
    G.tmap[t] = ST_NONE;
  } else if (f.nw) {
    uint32_t k = baseN + R.r[1];
    G.newList[k * 2] = t;
    G.newList[k * 2 + 1] = baseK + R.r[2] + k;  // slotList index if it gets a slot
  } else if (f.kp) {
    uint32_t nb = baseN + R.r[1];
    G.slotList[baseK + R.r[2] + min(nb, nalloc)] = G.tmap[t];
  }
}

// tileAlloc (wgsl-blocks.js:26-33), deterministic pop order
__global__ void k_tile_alloc(const DParams* __restrict__ PP, DGrid G) {
  uint32_t nn = G.ctl[3], top1 = G.ctl[13], nalloc = G.ctl[14];
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nn; k += gridDim.x * blockDim.x) {
    if (k < nalloc) {
      uint32_t s = G.fstack[top1 - 1u - k], t = G.newList[k * 2], idx = G.newList[k * 2 + 1];
      G.tmap[t] = s;
      G.slotTile[s] = t;
      G.slotList[idx] = s;
      G.newList[k * 2 + 1] = s;
    } else {
      G.newList[k * 2 + 1] = ST_NONE;
    }
  }
  (void)PP;
}

// initNew (wgsl-blocks.js:38-47): no pressure, collider lookup of each cell
__global__ void k_init_new(const DParams* __restrict__ PP, const DCols* __restrict__ C, DGrid G) {
  const DParams& P = *PP;
  uint32_t tot = G.ctl[3] * 512u;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += gridDim.x * blockDim.x) {
    uint32_t s = G.newList[(t >> 9) * 2 + 1];
    if (s == ST_NONE) continue;
    uint32_t local = t & 511u, tl = G.newList[(t >> 9) * 2], tt = tl / P.tx;
    int3 p = make_int3(int32_t((tl % P.tx) * 8u + (local & 7u)), int32_t((tt % P.ty) * 8u + ((local >> 3) & 7u)), int32_t((tt / P.ty) * 8u + (local >> 6)));
    uint32_t c = s * 512u + local;
    G.pres[c] = 0.0f;
    G.ctype[c] = uint32_t(d_colAt(P, C, G.sdf, d_cellCenter(P, p.x, p.y, p.z)) + 1) << 8;
    if (G.velG) {
      G.velG[c * 3] = 0.0f;
      G.velG[c * 3 + 1] = 0.0f;
      G.velG[c * 3 + 2] = 0.0f;
    }
  }
}

// ---------------------------------------------------------------------------
// clearSlots (wgsl-solver.js:5-15): every slot in use plus the spare slot 0
__global__ void k_clear(const DParams* __restrict__ PP, DGrid G) {
  uint32_t sc = G.ctl[4];
  uint32_t tot = (sc + 1u) * 512u;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    G.ctl[0] = 0u;
    G.ctl[1] = 0u;
  }
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += gridDim.x * blockDim.x) {
    uint32_t blk = t >> 9;
    uint32_t s = blk < sc ? G.slotList[blk] : 0u;
    uint32_t c = s * 512u + (t & 511u);
    G.cnt[c] = 0;
    G.accV[c * 3] = 0; G.accV[c * 3 + 1] = 0; G.accV[c * 3 + 2] = 0;
    G.accW[c * 3] = 0; G.accW[c * 3 + 1] = 0; G.accW[c * 3 + 2] = 0;
    if (s == 0u) G.rhs[c] = 8.0f;
  }
  (void)PP;
}

// ---------------------------------------------------------------------------
// p2g (wgsl-solver.js:16-40): fixed-point sums, so the result does not depend on
// the order the atomics land in.
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off));
  return v;
}
__global__ void __launch_bounds__(WG) k_p2g_global(const DParams* __restrict__ PP, DGrid G, const float4* __restrict__ pos, const float4* __restrict__ vel) {
  const DParams& P = *PP;
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  bool ok = i < P.n;
  float4 p4 = ok ? pos[i] : make_float4(0.f, -1.0e6f, 0.f, 0.f);
  ok = ok && !(p4.y < ST_DEADY);
  float3 p = make_float3(p4.x, p4.y, p4.z);
  int3 cc = d_particleCell(P, p);
  ok = ok && G.tmap[d_tileOf(P, cc.x, cc.y, cc.z)] != ST_NONE;
  float4 v4 = ok ? vel[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  float3 v = make_float3(v4.x, v4.y, v4.z);
  uint32_t bits = ok ? st_f2u(f3len(v)) : 0u;
  bits = wave_max_u32(bits);
  if (lane_id() == 0 && bits) atomicMax(&G.ctl[6], bits);
  if (!ok) return;
  for (uint32_t axis = 0; axis < 3; axis++) {
    float3 q = make_float3((p.x - P.x0) * P.ih, (p.y - P.y0) * P.ih, (p.z - P.z0) * P.ih);
    if (axis == 0) { q.y = q.y - 0.5f; q.z = q.z - 0.5f; }
    else if (axis == 1) { q.x = q.x - 0.5f; q.z = q.z - 0.5f; }
    else { q.x = q.x - 0.5f; q.y = q.y - 0.5f; }
    float3 f0 = make_float3(floorf(q.x), floorf(q.y), floorf(q.z));
    float3 fr = f3sub(q, f0);
    int3 i0 = make_int3(st_f2i_sat(f0.x), st_f2i_sat(f0.y), st_f2i_sat(f0.z));
    float vc = f3get(v, axis);
    for (uint32_t c = 0; c < 8; c++) {
      int32_t dx = int32_t(c & 1u), dy = int32_t((c >> 1) & 1u), dz = int32_t((c >> 2) & 1u);
      float w = (dx == 1 ? fr.x : 1.0f - fr.x) * (dy == 1 ? fr.y : 1.0f - fr.y) * (dz == 1 ? fr.z : 1.0f - fr.z);
      if (w <= 0.0f) continue;
      uint32_t idx = d_faceIndex(P, G.tmap, axis, i0.x + dx, i0.y + dy, i0.z + dz);
      atomicAdd(&G.accV[idx], st_f2i_sat(rintf(w * vc * P.scale)));
      atomicAdd(&G.accW[idx], st_f2i_sat(rintf(w * P.scale)));
    }
  }
  atomicAdd(&G.cnt[d_cellIdx(P, G.tmap, cc.x, cc.y, cc.z)], 1);
}

// ---------------------------------------------------------------------------
// per active cell kernels: t -> slot slotList[t >> 9], local t & 511 (acell,
// wgsl-core.js:96-103)
#define FOR_ACTIVE(G)                                                                   \
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, tot_ = (G).ctl[4] * 512u; t < tot_; \
       t += gridDim.x * blockDim.x)

// solidBuild (wgsl-solver.js:41-45)
__global__ void k_solid_build(const DParams* __restrict__ PP, const DCols* __restrict__ C, DGrid G) {
  const DParams& P = *PP;
  FOR_ACTIVE(G) {
    uint32_t s = G.slotList[t >> 9], local = t & 511u, c = s * 512u + local;
    int3 p = d_slotCoord(P, G.slotTile, s, local);
    G.ctype[c] = (G.ctype[c] & 0xffff00ffu) | (uint32_t(d_colAt(P, C, G.sdf, d_cellCenter(P, p.x, p.y, p.z)) + 1) << 8);
  }
}

__device__ __forceinline__ float3 d_velGVec(const DParams& P, const DGrid& G, float3 x);

// classify (wgsl-solver.js:46-74), plus the per-colour fluid counts of buildList
// (wgsl-solver.js:152-158) as statistics
__global__ void __launch_bounds__(WG) k_classify(const DParams* __restrict__ PP, DGrid G) {
  const DParams& P = *PP;
  uint32_t tot = G.ctl[4] * 512u;
  uint32_t stride = gridDim.x * blockDim.x;
  uint32_t nred = 0, nblack = 0;
  for (uint32_t t0 = blockIdx.x * blockDim.x; t0 < tot; t0 += stride) {
    uint32_t t = t0 + threadIdx.x;
    if (t >= tot) break;
    uint32_t s = G.slotList[t >> 9], local = t & 511u, c = s * 512u + local;
    int3 a = d_slotCoord(P, G.slotTile, s, local);
    uint32_t ty = 0;
    uint32_t sb = G.ctype[c] & 0xffffff00u;
    int32_t n = G.cnt[c];
    if (!d_inDom(P, a)) ty = 0;
    else if ((sb & 0xff00u) != 0u) ty = 2;
    else if (P.nb != 0u && d_phiOfBits(sb) < -1.0f &&
             (d_phiOfBits(sb) < -P.band || d_calmV(P, d_velGVec(P, G, d_cellCenter(P, a.x, a.y, a.z))))) ty = 1;
    else if (n > 0) {
      ty = 1;
      if (float(n) < P.sheetMin) {
        int32_t dth = st_f2i_sat(ceilf(0.6f * P.rest));
        bool dense = false;
        for (uint32_t q = 0; q < 6; q++) {
          int3 r = a;
          if (q == 0) r.x -= 1; else if (q == 1) r.x += 1; else if (q == 2) r.y -= 1; else if (q == 3) r.y += 1; else if (q == 4) r.z -= 1; else r.z += 1;
          if (r.x < 0 || r.y < 0 || r.z < 0 || r.x >= int32_t(P.nx) || r.y >= int32_t(P.ny) || r.z >= int32_t(P.nz)) continue;
          if (G.cnt[d_cellIdx(P, G.tmap, r.x, r.y, r.z)] >= dth) dense = true;
        }
        if (!dense) ty = 0;
      }
    }
    G.ctype[c] = ty | sb;
    if (ty != 1u) G.pres[c] = 0.0f;
    if (ty == 1u) {
      if (((a.x + a.y + a.z) & 1) == 0) nred++; else nblack++;
    }
  }
  // wave sums, then one atomic per wave and colour
  for (int off = 32; off > 0; off >>= 1) {
    nred += (uint32_t)__shfl_xor((int)nred, off);
    nblack += (uint32_t)__shfl_xor((int)nblack, off);
  }
  if (lane_id() == 0) {
    if (nred) atomicAdd(&G.ctl[0], nred);
    if (nblack) atomicAdd(&G.ctl[1], nblack);
  }
}

// normBody / normFaces (wgsl-solver.js:75-105)
__global__ void k_norm_faces(const DParams* __restrict__ PP, const DCols* __restrict__ C, DGrid G) {
  const DParams& P = *PP;
  FOR_ACTIVE(G) {
    uint32_t s = G.slotList[t >> 9], local = t & 511u, c = s * 512u + local;
    int3 at = d_slotCoord(P, G.slotTile, s, local);
    for (uint32_t axis = 0; axis < 3; axis++) {
      uint32_t f = c * 3u + axis;
      int3 lo = at;
      if (axis == 0) lo.x -= 1; else if (axis == 1) lo.y -= 1; else lo.z -= 1;
      uint32_t tl = d_ctypeOf(P, G.tmap, G.ctype, lo.x, lo.y, lo.z), th = d_ctypeOf(P, G.tmap, G.ctype, at.x, at.y, at.z);
      int32_t ws = G.accW[f], vs = G.accV[f];
      float v = ws > 0 ? float(vs) / float(max(ws, 1)) : 0.0f;
      float old = v;
      if (axis == 1) v = v - P.gravity * P.dt;
      if (tl == 2u || th == 2u) {
        float sv = 0.0f;
        int32_t ci = -1;
        if (tl == 2u) ci = d_colAtCell(P, G.tmap, G.ctype, lo.x, lo.y, lo.z);
        if (ci < 0 && th == 2u) ci = d_colAtCell(P, G.tmap, G.ctype, at.x, at.y, at.z);
        float3 fp = d_facePos(P, at, axis);
        if (ci >= 0) sv = f3get(d_colSurfVel(C, uint32_t(ci), fp), axis);
        v = sv;
        old = sv;
      }
      G.accW[f] = (int32_t)st_f2u(old);
      G.accV[f] = (int32_t)st_f2u(v);
    }
  }
}

__device__ __forceinline__ float gvR(const DGrid& G, uint32_t f) { return st_u2f((uint32_t)G.accV[f]); }
__device__ __forceinline__ float gv0R(const DGrid& G, uint32_t f) { return st_u2f((uint32_t)G.accW[f]); }

// SOR stencil word: bit 15 fluid, bits 0-2 non-solid neighbours, bits 4-9 which
// neighbours (x-, x+, y-, y+, z-, z+) are fluid.
#define STEN_FLUID 0x8000u

// divergence (wgsl-solver.js:106-123); also records the SOR stencil of the cell,
// which depends only on the cell types fixed by classify.
__global__ void k_divergence(const DParams* __restrict__ PP, DGrid G) {
  const DParams& P = *PP;
  FOR_ACTIVE(G) {
    uint32_t s = G.slotList[t >> 9], local = t & 511u, c = s * 512u + local;
    uint32_t ct = G.ctype[c];
    if ((ct & 0xffu) != 1u) {
      G.rhs[c] = 0.0f;
      G.sten[c] = 0u;
      continue;
    }
    int3 a = d_slotCoord(P, G.slotTile, s, local);
    int32_t i = a.x, j = a.y, k = a.z;
    float div = (gvR(G, d_faceIndex(P, G.tmap, 0, i + 1, j, k)) - gvR(G, d_faceIndex(P, G.tmap, 0, i, j, k)) +
                 gvR(G, d_faceIndex(P, G.tmap, 1, i, j + 1, k)) - gvR(G, d_faceIndex(P, G.tmap, 1, i, j, k)) +
                 gvR(G, d_faceIndex(P, G.tmap, 2, i, j, k + 1)) - gvR(G, d_faceIndex(P, G.tmap, 2, i, j, k))) * P.ih;
    uint32_t tys[6];
    tys[0] = d_ctypeOf(P, G.tmap, G.ctype, i - 1, j, k);
    tys[1] = d_ctypeOf(P, G.tmap, G.ctype, i + 1, j, k);
    tys[2] = d_ctypeOf(P, G.tmap, G.ctype, i, j - 1, k);
    tys[3] = d_ctypeOf(P, G.tmap, G.ctype, i, j + 1, k);
    tys[4] = d_ctypeOf(P, G.tmap, G.ctype, i, j, k - 1);
    tys[5] = d_ctypeOf(P, G.tmap, G.ctype, i, j, k + 1);
    bool interior = true;
    uint32_t nns = 0, mask = 0;
    for (int q = 0; q < 6; q++) {
      if (tys[q] == 0u) interior = false;
      if (tys[q] != 2u) nns++;
      if (tys[q] == 1u) mask |= 1u << q;
    }
    float dens = float(G.cnt[c]) / P.rest - 1.0f;
    float tgt = P.comp * (interior ? dens : fmaxf(dens, 0.0f));
    if (P.nb != 0u && d_phiOfBits(ct) < -1.0f) tgt = 0.0f;
    G.rhs[c] = (tgt - div) * P.h * P.h / P.dt;
    G.sten[c] = STEN_FLUID | nns | (mask << 4);
  }
}

// relax / rbRed / rbBlack (wgsl-solver.js:124-140, 161-164): one colour over the
// cells of every tile in use.  Cells of one colour read only the other colour,
// so this is the reference's iterate exactly.
__global__ void k_sor(const DParams* __restrict__ PP, DGrid G, uint32_t color) {
  const DParams& P = *PP;
  FOR_ACTIVE(G) {
    uint32_t s = G.slotList[t >> 9], local = t & 511u, c = s * 512u + local;
    uint32_t st = G.sten[c];
    if (!(st & STEN_FLUID)) continue;
    int3 p = d_slotCoord(P, G.slotTile, s, local);
    if (uint32_t((p.x + p.y + p.z) & 1) != color) continue;
    uint32_t nns = st & 7u, mask = (st >> 4) & 63u;
    if (nns == 0u) {
      G.pres[c] = 0.0f;
      continue;
    }
    float sum = 0.0f;
    if (mask & 1u) sum = sum + G.pres[d_cellIdx(P, G.tmap, p.x - 1, p.y, p.z)];
    if (mask & 2u) sum = sum + G.pres[d_cellIdx(P, G.tmap, p.x + 1, p.y, p.z)];
    if (mask & 4u) sum = sum + G.pres[d_cellIdx(P, G.tmap, p.x, p.y - 1, p.z)];
    if (mask & 8u) sum = sum + G.pres[d_cellIdx(P, G.tmap, p.x, p.y + 1, p.z)];
    if (mask & 16u) sum = sum + G.pres[d_cellIdx(P, G.tmap, p.x, p.y, p.z - 1)];
    if (mask & 32u) sum = sum + G.pres[d_cellIdx(P, G.tmap, p.x, p.y, p.z + 1)];
    float pg = (G.rhs[c] + sum) / float(nns);
    float pc = G.pres[c];
    G.pres[c] = pc + P.omega * (pg - pc);
  }
}

// projBody / project (wgsl-solver.js:165-180)
__global__ void k_project(const DParams* __restrict__ PP, DGrid G) {
  const DParams& P = *PP;
  FOR_ACTIVE(G) {
    uint32_t s = G.slotList[t >> 9], local = t & 511u, c = s * 512u + local;
    int3 at = d_slotCoord(P, G.slotTile, s, local);
    for (uint32_t axis = 0; axis < 3; axis++) {
      int3 lo = at;
      if (axis == 0) lo.x -= 1; else if (axis == 1) lo.y -= 1; else lo.z -= 1;
      uint32_t tl = d_ctypeOf(P, G.tmap, G.ctype, lo.x, lo.y, lo.z), th = d_ctypeOf(P, G.tmap, G.ctype, at.x, at.y, at.z);
      if (tl == 2u || th == 2u) continue;
      if (tl != 1u && th != 1u) continue;
      float pl = 0.0f, ph = 0.0f;
      if (tl == 1u) pl = G.pres[d_cellIdx(P, G.tmap, lo.x, lo.y, lo.z)];
      if (th == 1u) ph = G.pres[c];
      uint32_t f = c * 3u + axis;
      G.accV[f] = (int32_t)st_f2u(gvR(G, f) - P.dt * P.ih * (ph - pl));
    }
  }
}

// ---------------------------------------------------------------------------
// g2p (wgsl-solver.js:185-275)

__device__ __forceinline__ float3 d_collideP(const DParams& P, const DCols* __restrict__ C, const float* __restrict__ sdf, float3 pin) {
  float3 p = pin;
  float m = P.margin;
  for (uint32_t k = 0; k < P.ncols; k++) {
    if (!d_colNear(P, C, k, p)) continue;
    float phi;
    float3 wn;
    d_sdfEval(C, sdf, k, p, phi, wn);
    if (phi < P.q0) p = f3add(p, f3mul(wn, P.q0 - phi));
  }
  const float big = 1.0e9f;
  float3 lo = make_float3((P.open & 1u) ? -big : P.x0 + m, (P.open & 4u) ? -big : P.y0 + m, (P.open & 16u) ? -big : P.z0 + m);
  float3 hi = make_float3((P.open & 2u) ? big : P.x1 - m, (P.open & 8u) ? big : P.y1 - m, (P.open & 32u) ? big : P.z1 - m);
  return make_float3(st_clampf(p.x, lo.x, hi.x), st_clampf(p.y, lo.y, hi.y), st_clampf(p.z, lo.z, hi.z));
}
__device__ __forceinline__ float d_hsh3(float3 p) {
  float3 q = make_float3(st_fractf(p.x * 0.3183099f + 0.1f), st_fractf(p.y * 0.3183099f + 0.2f), st_fractf(p.z * 0.3183099f + 0.3f));
  q = f3mul(q, 17.0f);
  return st_fractf(q.x * q.y * q.z * (q.x + q.y + q.z));
}
__device__ __forceinline__ float d_vnz(float3 x) {
  float3 i = make_float3(floorf(x.x), floorf(x.y), floorf(x.z));
  float3 f = make_float3(st_fractf(x.x), st_fractf(x.y), st_fractf(x.z));
  float3 u = make_float3(f.x * f.x * (3.0f - 2.0f * f.x), f.y * f.y * (3.0f - 2.0f * f.y), f.z * f.z * (3.0f - 2.0f * f.z));
  float a = st_mixf(st_mixf(d_hsh3(i), d_hsh3(f3add(i, make_float3(1, 0, 0))), u.x), st_mixf(d_hsh3(f3add(i, make_float3(0, 1, 0))), d_hsh3(f3add(i, make_float3(1, 1, 0))), u.x), u.y);
  float b = st_mixf(st_mixf(d_hsh3(f3add(i, make_float3(0, 0, 1))), d_hsh3(f3add(i, make_float3(1, 0, 1))), u.x), st_mixf(d_hsh3(f3add(i, make_float3(0, 1, 1))), d_hsh3(f3add(i, make_float3(1, 1, 1))), u.x), u.y);
  return st_mixf(a, b, u.z);
}

__device__ __forceinline__ void d_sampleBoth(const DParams& P, const DGrid& G, float3 p, float3& vn, float3& vo) {
  float rn[3], ro[3];
  for (uint32_t axis = 0; axis < 3; axis++) {
    float3 g = make_float3((p.x - P.x0) * P.ih, (p.y - P.y0) * P.ih, (p.z - P.z0) * P.ih);
    if (axis == 0) { g.y = g.y - 0.5f; g.z = g.z - 0.5f; }
    else if (axis == 1) { g.x = g.x - 0.5f; g.z = g.z - 0.5f; }
    else { g.x = g.x - 0.5f; g.y = g.y - 0.5f; }
    float3 f0 = make_float3(floorf(g.x), floorf(g.y), floorf(g.z));
    float3 fr = f3sub(g, f0);
    int3 i0 = make_int3(st_f2i_sat(f0.x), st_f2i_sat(f0.y), st_f2i_sat(f0.z));
    float sn = 0.0f, so = 0.0f;
    for (uint32_t c = 0; c < 8; c++) {
      int32_t dx = int32_t(c & 1u), dy = int32_t((c >> 1) & 1u), dz = int32_t((c >> 2) & 1u);
      float w = (dx == 1 ? fr.x : 1.0f - fr.x) * (dy == 1 ? fr.y : 1.0f - fr.y) * (dz == 1 ? fr.z : 1.0f - fr.z);
      uint32_t f = d_faceIndex(P, G.tmap, axis, i0.x + dx, i0.y + dy, i0.z + dz);
      sn = sn + w * gvR(G, f);
      so = so + w * gv0R(G, f);
    }
    rn[axis] = sn;
    ro[axis] = so;
  }
  vn = make_float3(rn[0], rn[1], rn[2]);
  vo = make_float3(ro[0], ro[1], ro[2]);
}

__global__ void __launch_bounds__(WG) k_g2p(const DParams* __restrict__ PP, const DCols* __restrict__ C, DGrid G,
                                            const float4* __restrict__ posIn, const float4* __restrict__ velIn,
                                            float4* __restrict__ posOut, float4* __restrict__ velOut) {
  const DParams& P = *PP;
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  float4 p4 = posIn[i];
  float4 v04 = velIn[i];
  if (p4.y < ST_DEADY) {
    posOut[i] = p4;
    velOut[i] = v04;
    return;
  }
  float3 p = make_float3(p4.x, p4.y, p4.z);
  float3 v0 = make_float3(v04.x, v04.y, v04.z);
  int3 fcell = d_particleCell(P, p);
  float3 vn = make_float3(v0.x - 0.0f, v0.y - P.gravity * P.dt, v0.z - 0.0f);
  float3 vo = v0;
  if (G.tmap[d_tileOf(P, fcell.x, fcell.y, fcell.z)] != ST_NONE) d_sampleBoth(P, G, p, vn, vo);
  float3 fl = f3add(v0, f3sub(vn, vo));
  float3 v = make_float3(st_mixf(vn.x, fl.x, P.flip), st_mixf(vn.y, fl.y, P.flip), st_mixf(vn.z, fl.z, P.flip));
  int3 bc = make_int3(st_f2i_sat(floorf((p.x - P.x0) * P.ih - 0.5f)), st_f2i_sat(floorf((p.y - P.y0) * P.ih - 0.5f)), st_f2i_sat(floorf((p.z - P.z0) * P.ih - 0.5f)));
  float csum = 8.0f * P.rest;
  bool fast = P.q12 > 0.5f;
  if (P.q9 < 0.5f && fast) {
    csum = 8.0f * float(G.cnt[d_cellIdx(P, G.tmap, fcell.x, fcell.y, fcell.z)]);
  } else if (P.q9 < 0.5f) {
    csum = 0.0f;
    for (int32_t dz = 0; dz <= 1; dz++)
      for (int32_t dy = 0; dy <= 1; dy++)
        for (int32_t dx = 0; dx <= 1; dx++) {
          int32_t cx = d_iclamp(bc.x + dx, 0, int32_t(P.nx) - 1), cy = d_iclamp(bc.y + dy, 0, int32_t(P.ny) - 1), cz = d_iclamp(bc.z + dz, 0, int32_t(P.nz) - 1);
          csum = csum + float(G.cnt[d_cellIdx(P, G.tmap, cx, cy, cz)]);
        }
  }
  float rho = csum / (8.0f * P.rest);
  float air = 1.0f - st_smoothstepf(0.008f, 0.1f, rho);
  float shield = st_smoothstepf(0.1f, 0.65f, rho);
  float exposure = 1.0f - 0.97f * shield;
  float spd0 = f3len(v);
  float sens = fmaxf(P.q6, 0.05f);
  float pacc = f3len(f3sub(vn, vo)) / P.dt;
  float settle = st_smoothstepf(0.1f, 0.5f, P.q8);
  float impact = st_clampf((pacc / sens - 15.0f) / 80.0f, 0.0f, 1.0f) * settle;
  float turb = st_clampf((f3len(f3sub(v0, vo)) / sens - 0.35f) / 1.5f, 0.0f, 1.0f) * settle;
  float spray = air * (0.25f + 0.75f * st_clampf(spd0 / 4.0f, 0.0f, 1.0f));
  float tau = fmaxf(P.q5, 0.05f) * st_mixf(1.0f, 0.35f, st_smoothstepf(0.7f, 1.0f, rho)) * st_mixf(1.0f, 2.5f, air);
  float foam = v04.w * st_expf(-P.dt / tau);
  foam = fmaxf(foam, fmaxf(spray * 0.9f, fmaxf(impact, turb * 0.8f)));
  float3 gust = make_float3(0.0f, 0.0f, 0.0f);
  if (P.q7 > 0.0f && P.q4 > 0.0f && (!fast || shield < 0.9f)) {
    float3 gp = f3mul(p, 0.55f);
    float tt = P.q8 * 0.6f;
    gust = f3sub(f3mul(make_float3(d_vnz(f3add(gp, make_float3(tt, 0.0f, 0.0f))), d_vnz(f3add(gp, make_float3(17.3f, tt, 4.1f))),
                                   d_vnz(f3add(gp, make_float3(5.7f, 9.2f, tt)))), 2.0f),
                 make_float3(1.0f, 1.0f, 1.0f));
  }
  float kd = P.q4 * (0.04f + 1.5f * exposure * exposure);
  float3 airv = f3add(make_float3(P.q1, P.q2, P.q3), f3mul(gust, P.q7));
  v = f3add(v, f3mul(f3sub(airv, v), 1.0f - st_expf(-kd * P.dt)));
  float sp = f3len(v);
  if (sp > P.vmax) v = f3mul(v, P.vmax / sp);
  float3 pt = f3add(p, f3mul(v, P.dt));
  float3 pc = d_collideP(P, C, G.sdf, pt);
  float3 dsp = f3sub(pc, pt);
  if (f3dot(dsp, dsp) > 1.0e-12f) {
    float3 nrm = f3div(dsp, f3len(dsp));
    float vd = f3dot(v, nrm);
    if (vd < 0.0f) v = f3sub(v, f3mul(nrm, vd));
  }
  uint32_t o = P.open;
  if (o != 0u && (((o & 1u) && pc.x < P.x0) || ((o & 2u) && pc.x > P.x1) || ((o & 4u) && pc.y < P.y0) || ((o & 8u) && pc.y > P.y1) ||
                  ((o & 16u) && pc.z < P.z0) || ((o & 32u) && pc.z > P.z1))) {
    posOut[i] = make_float4(0.0f, -1.0e6f, 0.0f, 0.0f);
    velOut[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  posOut[i] = make_float4(pc.x, pc.y, pc.z, fminf(f3len(v), 63.0f) + 64.0f * floorf(foam * 15.0f + 0.5f));
  velOut[i] = make_float4(v.x, v.y, v.z, foam);
}

// kept grid velocity at a point, all three components (wgsl-core.js:185-197)
__device__ __forceinline__ float d_velGAxis(const DParams& P, const DGrid& G, float3 x, uint32_t axis) {
  float3 g = make_float3((x.x - P.x0) * P.ih, (x.y - P.y0) * P.ih, (x.z - P.z0) * P.ih);
  if (axis == 0) { g.y = g.y - 0.5f; g.z = g.z - 0.5f; }
  else if (axis == 1) { g.x = g.x - 0.5f; g.z = g.z - 0.5f; }
  else { g.x = g.x - 0.5f; g.y = g.y - 0.5f; }
  float3 f0 = make_float3(floorf(g.x), floorf(g.y), floorf(g.z));
  float3 fr = f3sub(g, f0);
  int3 i0 = make_int3(st_f2i_sat(f0.x), st_f2i_sat(f0.y), st_f2i_sat(f0.z));
  float s = 0.0f;
  for (uint32_t c = 0; c < 8; c++) {
    int32_t dx = int32_t(c & 1u), dy = int32_t((c >> 1) & 1u), dz = int32_t((c >> 2) & 1u);
    float w = (dx == 1 ? fr.x : 1.0f - fr.x) * (dy == 1 ? fr.y : 1.0f - fr.y) * (dz == 1 ? fr.z : 1.0f - fr.z);
    s = s + w * G.velG[d_faceIndex(P, G.tmap, axis, i0.x + dx, i0.y + dy, i0.z + dz)];
  }
  return s;
}
__device__ __forceinline__ float3 d_velGVec(const DParams& P, const DGrid& G, float3 x) {
  return make_float3(d_velGAxis(P, G, x, 0), d_velGAxis(P, G, x, 1), d_velGAxis(P, G, x, 2));
}

// ---------------------------------------------------------------------------
// readback in id order (unsort, wgsl-blocks.js:98-99)
__global__ void k_unsort(const DParams* __restrict__ PP, const float4* __restrict__ pos, const uint32_t* __restrict__ pid, float4* __restrict__ out, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[pid[i]] = pos[i];
  (void)PP;
}

// ---------------------------------------------------------------------------
// launch wrappers

static inline uint32_t gridFor(size_t items, uint32_t block) { return uint32_t((items + block - 1) / block); }

void launch_fill_u32(hipStream_t st, uint32_t* p, size_t n, uint32_t v) {
  if (!n) return;
  uint32_t g = std::min<uint32_t>(gridFor(n, WG), 8192u);
  hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(WG), 0, st, p, n, v);
}

void launch_blocks(hipStream_t st, const LaunchCtx& L) {
  uint32_t nt = L.ntiles;
  launch_fill_u32(st, L.G.need, nt, 0u);
  if (L.n) hipLaunchKernelGGL(k_tile_mark, dim3(gridFor(L.n, WG)), dim3(WG), 0, st, L.P, L.G, (const float4*)L.G.pos);
  uint32_t nwg = gridFor(nt, WG);
  hipLaunchKernelGGL(k_tile_count, dim3(nwg), dim3(WG), 0, st, L.P, L.G);
  hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, st, L.P, L.G, nwg);
  hipLaunchKernelGGL(k_tile_scatter, dim3(nwg), dim3(WG), 0, st, L.P, L.G);
  hipLaunchKernelGGL(k_tile_alloc, dim3(std::max<uint32_t>(1u, std::min<uint32_t>(gridFor(nt, WG), 256u))), dim3(WG), 0, st, L.P, L.G);
  hipLaunchKernelGGL(k_init_new, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.C, L.G);
}

void launch_clear(hipStream_t st, const LaunchCtx& L) {
  hipLaunchKernelGGL(k_clear, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.G);
}

void launch_p2g(hipStream_t st, const LaunchCtx& L) {
  if (!L.n) return;
  hipLaunchKernelGGL(k_p2g_global, dim3(gridFor(L.n, WG)), dim3(WG), 0, st, L.P, L.G, (const float4*)L.G.pos, (const float4*)L.G.vel);
}

void launch_classify(hipStream_t st, const LaunchCtx& L, bool solid) {
  if (solid) hipLaunchKernelGGL(k_solid_build, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.C, L.G);
  hipLaunchKernelGGL(k_classify, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.G);
}

void launch_norm_faces(hipStream_t st, const LaunchCtx& L) {
  hipLaunchKernelGGL(k_norm_faces, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.C, L.G);
}

void launch_divergence(hipStream_t st, const LaunchCtx& L) {
  hipLaunchKernelGGL(k_divergence, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.G);
}

void launch_pressure(hipStream_t st, const LaunchCtx& L, int iters) {
  for (int it = 0; it < iters; it++) {
    hipLaunchKernelGGL(k_sor, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.G, 0u);
    hipLaunchKernelGGL(k_sor, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.G, 1u);
  }
}

void launch_project(hipStream_t st, const LaunchCtx& L) {
  hipLaunchKernelGGL(k_project, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.G);
}

void launch_g2p(hipStream_t st, const LaunchCtx& L) {
  if (!L.n) return;
  hipLaunchKernelGGL(k_g2p, dim3(gridFor(L.n, WG)), dim3(WG), 0, st, L.P, L.C, L.G, (const float4*)L.G.pos, (const float4*)L.G.vel,
                     L.G.pos, L.G.vel);
}

void launch_unsort(hipStream_t st, const LaunchCtx& L, float4* out, uint32_t n) {
  if (!n) return;
  hipLaunchKernelGGL(k_unsort, dim3(gridFor(n, WG)), dim3(WG), 0, st, L.P, (const float4*)L.G.pos, (const uint32_t*)L.G.pid, out, n);
}

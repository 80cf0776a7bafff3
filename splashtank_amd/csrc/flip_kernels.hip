// flip_kernels.hip - HIP kernels of the FLIP substep for gfx950 (wave64).
//
// Layout of the work:
//  * particles are binned by the tile of their cell every substep with a
//    counting sort of indices (binIdx); g2p bins the positions it writes for
//    the next substep, so p2g and g2p run one workgroup per tile over a
//    contiguous run of bin entries;
//  * p2g accumulates the tile's 10^3-cell window (the tile plus one cell all
//    round) with LDS integer atomics and adds the window to HBM once; the sums
//    are fixed point, so the order the atomics land in does not matter;
//  * g2p stages the same window of face velocities and counts in LDS;
//  * cell kernels are grid-stride loops over the device-side count of tiles in
//    use (ctl[4]), so no launch needs a host readback and all can live in a graph;
//  * the red-black SOR runs every iteration of a solve in one cooperative
//    launch, each CU keeping the pressure of a Morton-ordered range of tiles in
//    LDS and exchanging face halos with the others through a grid barrier; a
//    per-iteration tile-window kernel is the other path.
// The reference's kernels each of these replaces are cited at the kernel.
#include "flip_device.h"
#include "flip_kernels.h"
#include "grid_barrier.h"

#include <algorithm>
#include <utility>
#include <vector>

#define WG 256
#define WIN 10      // window edge: tile of 8 plus one cell each side
#define WIN3 1000
#define SW3_1 1728  // 12^3: pressure window of one SOR iteration
#define SOR_NRED 448  // red cells one SOR tile relaxes: its own 256 and 6 x 32 across its faces

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off));
  return v;
}
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
  for (int off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off);
  return v;
}

// ---------------------------------------------------------------------------
// generic fills (no hipMemsetAsync, so the launches can live in a graph)

__global__ void k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------------------
// blocks (wgsl-blocks.js:4-60) and the particle bins

// tileMark + markAround (wgsl-blocks.js:10-14, wgsl-core.js:239-246): every tile
// within two cells of a live particle is needed.  The same pass counts each
// particle into the bin of its cell's tile (dead particles into bin ntiles) and
// records the bin in binKey and its place within the bin in binRank.
// Consecutive particles mostly share a bin (the array is in bin order), so the
// counting is one atomic per run of equal bins in a wave.  k_mark_bin runs this
// over A at the start of a substep; g2p runs it as it writes the new state, for
// the next substep, when nothing else moves particles in between.
// Particle i at position q (valid lanes only).  Every lane of the wave must
// call this (lanes 0..26 do the marking).
__device__ __forceinline__ void bin_mark(const DParams& P, const DGrid& G, uint32_t i, bool valid, float4 q) {
  bool live = valid && !(q.y < ST_DEADY);
  uint32_t key = P.ntiles;
  // tiles [cc-2, cc+2] >> 3 relative to the particle's own tile, as bits
  // (dx+1) + 3 (dy+1) + 9 (dz+1); the span is at most two tiles per axis
  uint32_t nmask = 0;
  int32_t ox = 0, oy = 0, oz = 0;
  if (live) {
    int3 cc = d_particleCell(P, make_float3(q.x, q.y, q.z));
    key = d_tileOf(P, cc.x, cc.y, cc.z);
    ox = cc.x >> 3; oy = cc.y >> 3; oz = cc.z >> 3;
    // per axis, bits of the offsets -1, 0, +1 that the span covers
    uint32_t bx = 2u | (cc.x - 2 >= 0 && ((cc.x - 2) >> 3) < ox ? 1u : 0u) |
                  (cc.x + 2 <= int32_t(P.tx) * 8 - 1 && ((cc.x + 2) >> 3) > ox ? 4u : 0u);
    uint32_t by = 2u | (cc.y - 2 >= 0 && ((cc.y - 2) >> 3) < oy ? 1u : 0u) |
                  (cc.y + 2 <= int32_t(P.ty) * 8 - 1 && ((cc.y + 2) >> 3) > oy ? 4u : 0u);
    uint32_t bz = 2u | (cc.z - 2 >= 0 && ((cc.z - 2) >> 3) < oz ? 1u : 0u) |
                  (cc.z + 2 <= int32_t(P.tz) * 8 - 1 && ((cc.z + 2) >> 3) > oz ? 4u : 0u);
    uint32_t xy = ((by & 1u) ? bx : 0u) | ((by & 2u) ? bx << 3 : 0u) | ((by & 4u) ? bx << 6 : 0u);
    nmask = ((bz & 1u) ? xy : 0u) | ((bz & 2u) ? xy << 9 : 0u) | ((bz & 4u) ? xy << 18 : 0u);
  }
  // runs of equal bins in the wave: one bin count atomic per run (its result is
  // not waited for); lane l < 27 marks neighbour l of the run's tile if any
  // particle of the run needs it
  const uint32_t ln = lane_id();
  const int32_t mdx = int32_t(ln % 3u) - 1, mdy = int32_t((ln / 3u) % 3u) - 1, mdz = int32_t(ln / 9u) - 1;
  uint32_t rank = 0u;
  uint64_t rem = __ballot(valid);
  while (rem) {
    int leader = __ffsll((unsigned long long)rem) - 1;
    uint32_t k = (uint32_t)__shfl((int)key, leader);
    uint64_t m = __ballot(valid && key == k) & rem;
    bool inrun = (m >> ln) & 1ull;
    uint32_t um = inrun ? nmask : 0u;
    for (int off = 32; off > 0; off >>= 1) um |= (uint32_t)__shfl_xor((int)um, off);
    // the run's places in the bin: one returning atomic, read at the end
    uint32_t base = 0u;
    if (int(ln) == leader) base = atomicAdd(&G.binCount[k], (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, leader);
    if (inrun) rank = base + (uint32_t)__popcll(m & lanemask_lt());
    if (k < P.ntiles) {
      int32_t tx0 = __shfl(ox, leader), ty0 = __shfl(oy, leader), tz0 = __shfl(oz, leader);
      if (ln < 27u && ((um >> ln) & 1u))
        G.need[uint32_t(tx0 + mdx) + P.tx * (uint32_t(ty0 + mdy) + P.ty * uint32_t(tz0 + mdz))] = 1u;
    }
    rem &= ~m;
  }
  if (valid) {
    G.binKey[i] = key;
    G.binRank[i] = rank;
  }
}

__device__ __forceinline__ void mark_bin_body(const DParams* __restrict__ PP, const DGrid& G, uint32_t bid, uint32_t nbid) {
  const DParams& P = *PP;
  uint32_t i = bid * blockDim.x + threadIdx.x;
  bool valid = i < P.n;
  float4 q = valid ? G.pos[i] : make_float4(0.f, -1.0e6f, 0.f, 0.f);
  bin_mark(P, G, i, valid, q);
}
__global__ void __launch_bounds__(WG) k_mark_bin(const DParams* __restrict__ PP, DGrid G) {
  mark_bin_body(PP, G, blockIdx.x, gridDim.x);
}

// exclusive scan of n counts into out[0..n], out[n] = total; three launches
__device__ __forceinline__ void scan_blocks_body(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t n,
                                                 uint32_t* __restrict__ sums, uint32_t bid, uint32_t nbid) {
  __shared__ uint32_t wsum[WG / 64];
  uint32_t base = bid * (WG * 4) + threadIdx.x * 4;
  uint32_t v[4], s = 0;
  for (int q = 0; q < 4; q++) {
    v[q] = base + q < n ? in[base + q] : 0u;
    s += v[q];
  }
  uint32_t inc = s;
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t o = (uint32_t)__shfl_up((int)inc, off);
    if (int(lane_id()) >= off) inc += o;
  }
  uint32_t w = threadIdx.x >> 6;
  if (lane_id() == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t before = 0, tot = 0;
  for (uint32_t u = 0; u < WG / 64; u++) {
    if (u < w) before += wsum[u];
    tot += wsum[u];
  }
  uint32_t run = before + inc - s;
  for (int q = 0; q < 4; q++) {
    if (base + q < n) out[base + q] = run;
    run += v[q];
  }
  if (threadIdx.x == 0) sums[bid] = tot;
}
__global__ void __launch_bounds__(WG) k_scan_blocks(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t n,
                                                    uint32_t* __restrict__ sums) {
  scan_blocks_body(in, out, n, sums, blockIdx.x, gridDim.x);
}
__device__ __forceinline__ void scan_top_body(uint32_t* __restrict__ sums, uint32_t nb, uint32_t* __restrict__ out,
                                              uint32_t n, uint32_t bid, uint32_t nbid) {
  __shared__ uint32_t part[1024 / 64];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t b0 = 0; b0 < nb; b0 += 1024) {
    uint32_t i = b0 + threadIdx.x;
    uint32_t v = i < nb ? sums[i] : 0u, inc = v;
    for (int off = 1; off < 64; off <<= 1) {
      uint32_t o = (uint32_t)__shfl_up((int)inc, off);
      if (int(lane_id()) >= off) inc += o;
    }
    uint32_t w = threadIdx.x >> 6;
    if (lane_id() == 63) part[w] = inc;
    __syncthreads();
    uint32_t before = 0, tot = 0;
    for (uint32_t u = 0; u < 1024 / 64; u++) {
      if (u < w) before += part[u];
      tot += part[u];
    }
    if (i < nb) sums[i] = carry + before + inc - v;
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) out[n] = carry;
}
__global__ void __launch_bounds__(1024) k_scan_top(uint32_t* __restrict__ sums, uint32_t nb, uint32_t* __restrict__ out,
                                                   uint32_t n) {
  scan_top_body(sums, nb, out, n, blockIdx.x, gridDim.x);
}
__device__ __forceinline__ void scan_add_body(uint32_t* __restrict__ out, uint32_t n, const uint32_t* __restrict__ sums,
                                              uint32_t* __restrict__ copy, uint32_t bid, uint32_t nbid) {
  uint32_t base = bid * (WG * 4) + threadIdx.x * 4;
  uint32_t add = sums[bid];
  for (int q = 0; q < 4; q++)
    if (base + q < n) {
      uint32_t v = out[base + q] + add;
      out[base + q] = v;
      if (copy) copy[base + q] = v;
    }
}
__global__ void __launch_bounds__(WG) k_scan_add(uint32_t* __restrict__ out, uint32_t n, const uint32_t* __restrict__ sums,
                                                 uint32_t* __restrict__ copy) {
  scan_add_body(out, n, sums, copy, blockIdx.x, gridDim.x);
}

// bin order as an index list: binIdx[binStart[bin] + rank] = i.  The particles
// themselves stay where they are; p2g and g2p read them through the list (runs
// of a bin are mostly contiguous in A, which is in last step's bin order).
__device__ __forceinline__ void bin_scatter_body(const DParams* __restrict__ PP, const DGrid& G, uint32_t bid, uint32_t nbid) {
  uint32_t i = bid * blockDim.x + threadIdx.x;
  if (i >= PP->n) return;
  G.binIdx[G.binStart[G.binKey[i]] + G.binRank[i]] = i;
}
__global__ void __launch_bounds__(WG) k_bin_scatter(const DParams* __restrict__ PP, DGrid G) {
  bin_scatter_body(PP, G, blockIdx.x, gridDim.x);
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

__device__ __forceinline__ void tile_count_body(const DParams* __restrict__ PP, const DGrid& G, uint32_t bid, uint32_t nbid) {
  uint32_t t = bid * WG + threadIdx.x;
  TileFlags f = tile_flags(G, t, PP->ntiles);
  Ranks3 R = wg_ranks3(f.fr, f.nw, f.kp);
  if (threadIdx.x == 0) {
    G.wg[bid * 3 + 0] = R.tot[0];
    G.wg[bid * 3 + 1] = R.tot[1];
    G.wg[bid * 3 + 2] = R.tot[2];
  }
}
__global__ void __launch_bounds__(WG) k_tile_count(const DParams* __restrict__ PP, DGrid G) {
  tile_count_body(PP, G, blockIdx.x, gridDim.x);
}

// One workgroup: exclusive scan of the per-workgroup counts, then the stack
// arithmetic of tileFree/tileAlloc/tileFix (wgsl-blocks.js:16-35) in closed form:
// the k-th new tile (tile order) pops stack entry top1-1-k, so the first
// min(new, top1) new tiles get slots and the rest count as no room.
__device__ __forceinline__ void tile_scan_body(const DGrid& G, uint32_t nwg, uint32_t bid, uint32_t nbid) {
  __shared__ uint32_t part[3][1024 / 64];
  __shared__ uint32_t carry[3];
  if (threadIdx.x < 3) carry[threadIdx.x] = 0;
  __syncthreads();
  for (uint32_t base = 0; base < nwg; base += 1024) {
    uint32_t i = base + threadIdx.x;
    uint32_t v[3];
    for (int q = 0; q < 3; q++) v[q] = i < nwg ? G.wg[i * 3 + q] : 0u;
    uint32_t s[3] = {v[0], v[1], v[2]};
    for (int off = 1; off < 64; off <<= 1)
      for (int q = 0; q < 3; q++) {
        uint32_t o = (uint32_t)__shfl_up((int)s[q], off);
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
}
__global__ void __launch_bounds__(1024) k_tile_scan(DGrid G, uint32_t nwg) {
  tile_scan_body(G, nwg, blockIdx.x, gridDim.x);
}

// Freed tiles push their slot (tile order), new tiles are listed (tile order),
// kept tiles take their place in slotList (tile order).
__device__ __forceinline__ void tile_scatter_body(const DParams* __restrict__ PP, const DGrid& G, uint32_t bid, uint32_t nbid) {
  uint32_t t = bid * WG + threadIdx.x;
  TileFlags f = tile_flags(G, t, PP->ntiles);
  Ranks3 R = wg_ranks3(f.fr, f.nw, f.kp);
  uint32_t baseF = G.wg[bid * 3 + 0], baseN = G.wg[bid * 3 + 1], baseK = G.wg[bid * 3 + 2];
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
__global__ void __launch_bounds__(WG) k_tile_scatter(const DParams* __restrict__ PP, DGrid G) {
  tile_scatter_body(PP, G, blockIdx.x, gridDim.x);
}

// tileAlloc (wgsl-blocks.js:26-33), deterministic pop order
__device__ __forceinline__ void tile_alloc_body(const DGrid& G, uint32_t bid, uint32_t nbid) {
  uint32_t nn = G.ctl[3], top1 = G.ctl[13], nalloc = G.ctl[14];
  for (uint32_t k = bid * blockDim.x + threadIdx.x; k < nn; k += nbid * blockDim.x) {
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
}
__global__ void k_tile_alloc(DGrid G) {
  tile_alloc_body(G, blockIdx.x, gridDim.x);
}

// Slots of the 27 tiles around every tile in use (itself included), index
// (dx+1) + 3 (dy+1) + 9 (dz+1): 0 for a tile without storage (cellIdx maps it to
// the spare slot), WSLOT_OUT outside the tile grid.  Built once per substep after
// allocation so the window kernels need no tile-map lookups.
#define WSLOT_OUT 0xfffffffeu
__device__ __forceinline__ void nbr_body(const DParams* __restrict__ PP, const DGrid& G, uint32_t bid, uint32_t nbid) {
  const DParams& P = *PP;
  uint32_t tot = G.ctl[4] * 27u;
  for (uint32_t e = bid * blockDim.x + threadIdx.x; e < tot; e += nbid * blockDim.x) {
    uint32_t b = e / 27u, q = e - b * 27u;
    uint32_t s = G.slotList[b], t = G.slotTile[s];
    int32_t x = int32_t(t % P.tx) + int32_t(q % 3u) - 1, y = int32_t((t / P.tx) % P.ty) + int32_t((q / 3u) % 3u) - 1,
            z = int32_t(t / (P.tx * P.ty)) + int32_t(q / 9u) - 1;
    uint32_t v = WSLOT_OUT;
    if (x >= 0 && y >= 0 && z >= 0 && x < int32_t(P.tx) && y < int32_t(P.ty) && z < int32_t(P.tz)) {
      uint32_t m = G.tmap[uint32_t(x) + P.tx * (uint32_t(y) + P.ty * uint32_t(z))];
      v = m == ST_NONE ? 0u : m;
    }
    G.nbr[s * 27u + q] = v;
    G.nbrL[e] = v;
  }
}
__global__ void k_nbr(const DParams* __restrict__ PP, DGrid G) {
  nbr_body(PP, G, blockIdx.x, gridDim.x);
}

// initNew (wgsl-blocks.js:38-47): no pressure, collider lookup of each cell
__device__ __forceinline__ void init_new_body(const DParams* __restrict__ PP, const DCols* __restrict__ C, const DGrid& G,
                                              uint32_t bid, uint32_t nbid) {
  const DParams& P = *PP;
  uint32_t tot = G.ctl[3] * 512u;
  for (uint32_t t = bid * blockDim.x + threadIdx.x; t < tot; t += nbid * blockDim.x) {
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
__global__ void k_init_new(const DParams* __restrict__ PP, const DCols* __restrict__ C, DGrid G) {
  init_new_body(PP, C, G, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// clearSlots (wgsl-solver.js:5-15): the counts and face sums of the slots in use
// and of the spare slot 0.  p2g stores the counts and the faces of every slot
// in use (each tile's own cells), so only the spare, which p2g adds to with
// atomics, is cleared here.
__device__ __forceinline__ void clear_body(const DGrid& G, uint32_t bid, uint32_t nbid) {
  if (bid == 0 && threadIdx.x == 0) {
    G.ctl[0] = 0u;
    G.ctl[1] = 0u;
  }
  for (uint32_t t = bid * blockDim.x + threadIdx.x; t < 512u; t += nbid * blockDim.x) {
    const uint32_t c = t;  // slot 0
    G.cnt[c] = 0;
    G.accV[c * 3] = 0; G.accV[c * 3 + 1] = 0; G.accV[c * 3 + 2] = 0;
    G.accW[c * 3] = 0; G.accW[c * 3 + 1] = 0; G.accW[c * 3 + 2] = 0;
    G.rhs[c] = 8.0f;
  }
}
__global__ void k_clear(DGrid G) {
  clear_body(G, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// Tile windows.  The window of tile (ti, tj, tk) is cells [8ti-1, 8ti+8] per axis.
// A window cell's tile is one of the 27 around the tile: offset -1 at window
// index 0, +1 at 9, else 0.  wslots[27] holds their slots (0 when the tile has
// no storage, as cellIdx maps it; WSLOT_OUT when outside the tile grid).
__device__ __forceinline__ void load_wslots(const DParams& P, const DGrid& G, uint32_t s, uint32_t t, uint32_t* wslots, int3& org) {
  int32_t ti = int32_t(t % P.tx), tj = int32_t((t / P.tx) % P.ty), tk = int32_t(t / (P.tx * P.ty));
  org = make_int3(ti * 8 - 1, tj * 8 - 1, tk * 8 - 1);
  if (threadIdx.x < 27) wslots[threadIdx.x] = G.nbr[s * 27u + threadIdx.x];
}
__device__ __forceinline__ int32_t woff(int32_t w) { return w == 0 ? 0 : (w == WIN - 1 ? 2 : 1); }
// storage index of window cell (wx, wy, wz), or ST_NONE outside the tile grid
__device__ __forceinline__ uint32_t wcell(const uint32_t* wslots, int32_t wx, int32_t wy, int32_t wz, int3 org) {
  uint32_t s = wslots[woff(wx) + 3 * (woff(wy) + 3 * woff(wz))];
  if (s == WSLOT_OUT) return ST_NONE;
  int32_t gx = org.x + wx, gy = org.y + wy, gz = org.z + wz;
  return s * 512u + uint32_t(gx & 7) + 8u * (uint32_t(gy & 7) + 8u * uint32_t(gz & 7));
}
// window cell index wx + WIN (wy + WIN wz) of offsets in [0, WIN): 24-bit
// multiplies (full rate; a 32-bit multiply is a quarter-rate instruction)
__device__ __forceinline__ uint32_t win_idx(int32_t wx, int32_t wy, int32_t wz) {
  return uint32_t(wx) + __umul24(uint32_t(wy), uint32_t(WIN)) + __umul24(uint32_t(wz), uint32_t(WIN * WIN));
}
// index 0 .. P2G_HALO_CELLS-1 of window cell (wx, wy, wz) of the halo shell (a
// coordinate at 0 or WIN-1): the two full z planes, then per inner plane the
// two full y rows and the x ends of the rows between
__device__ __forceinline__ uint32_t shell_idx(uint32_t wx, uint32_t wy, uint32_t wz) {
  if (wz == 0u || wz == WIN - 1u) return (wz == 0u ? 0u : WIN * WIN) + wx + WIN * wy;
  const uint32_t ring = (wy == 0u || wy == WIN - 1u) ? (wy == 0u ? 0u : WIN) + wx : 2u * WIN + 2u * (wy - 1u) + (wx == 0u ? 0u : 1u);
  return 2u * WIN * WIN + (wz - 1u) * (4u * WIN - 4u) + ring;
}
static_assert(2 * WIN * WIN + (WIN - 2) * (4 * WIN - 4) == P2G_HALO_CELLS, "the halo shell's cells");

// window offset, clamped into the window: a guard for LDS only -- every face or
// cell a particle of the tile touches already lies inside (faceIndex's clamp moves
// an index toward the particle's own clamped cell, which is in the tile)
__device__ __forceinline__ int32_t wclamp(int32_t w) { return min(max(w, 0), WIN - 1); }
// the two clamped window offsets along one axis of a trilinear stencil at i0
// (faceIndex, wgsl-core.js:84-89): own axis clamps to [0, n], others to [0, n-1]
__device__ __forceinline__ void wpair(int32_t i0, int32_t hi, int32_t org, int32_t* w) {
  w[0] = wclamp(d_iclamp(i0, 0, hi) - org);
  w[1] = wclamp(d_iclamp(i0 + 1, 0, hi) - org);
}
// One dimension of a trilinear stencil over a window: the fraction of the
// coordinate q, its complement and the two clamped window offsets.  Per
// dimension the three face stencils of a point use two coordinates, the
// unshifted one (a face's own axis, clamped to [0, n]) and the one shifted by
// half a cell (the other two axes, clamped to [0, n-1]), so p2g and g2p work
// these out once for each of the six and pick them per axis.
struct WinAxis {
  float f, g;    // fraction and 1 - fraction
  int32_t w[2];  // the two clamped window offsets
};
__device__ __forceinline__ WinAxis win_axis(float q, int32_t hi, int32_t org) {
  WinAxis a;
  const float f0 = floorf(q);
  a.f = q - f0;
  a.g = 1.0f - a.f;
  wpair(st_f2i_sat(f0), hi, org, a.w);
  return a;
}
struct WinAxes {
  WinAxis xu, xs, yu, ys, zu, zs;  // unshifted and shifted, per dimension
  __device__ const WinAxis& x(uint32_t axis) const { return axis == 0 ? xu : xs; }
  __device__ const WinAxis& y(uint32_t axis) const { return axis == 1 ? yu : ys; }
  __device__ const WinAxis& z(uint32_t axis) const { return axis == 2 ? zu : zs; }
};
__device__ __forceinline__ WinAxes win_axes(const DParams& P, float3 p, int3 org) {
  const float ux = (p.x - P.x0) * P.ih, uy = (p.y - P.y0) * P.ih, uz = (p.z - P.z0) * P.ih;
  WinAxes a;
  a.xu = win_axis(ux, int32_t(P.nx), org.x);
  a.xs = win_axis(ux - 0.5f, int32_t(P.nx) - 1, org.x);
  a.yu = win_axis(uy, int32_t(P.ny), org.y);
  a.ys = win_axis(uy - 0.5f, int32_t(P.ny) - 1, org.y);
  a.zu = win_axis(uz, int32_t(P.nz), org.z);
  a.zs = win_axis(uz - 0.5f, int32_t(P.nz) - 1, org.z);
  return a;
}
// face (axis, i, j, k) clamped as faceIndex does (wgsl-core.js:84-89), as a window index or -1
__device__ __forceinline__ int32_t wface(const DParams& P, uint32_t axis, int32_t i, int32_t j, int32_t k, int3 org) {
  int32_t X = int32_t(P.nx), Y = int32_t(P.ny), Z = int32_t(P.nz);
  int32_t ci = d_iclamp(i, 0, axis == 0 ? X : X - 1), cj = d_iclamp(j, 0, axis == 1 ? Y : Y - 1), ck = d_iclamp(k, 0, axis == 2 ? Z : Z - 1);
  int32_t wx = ci - org.x, wy = cj - org.y, wz = ck - org.z;
  if (uint32_t(wx) >= uint32_t(WIN) || uint32_t(wy) >= uint32_t(WIN) || uint32_t(wz) >= uint32_t(WIN)) return -1;
  return wx + WIN * (wy + WIN * wz);
}

// ---------------------------------------------------------------------------
// p2g (wgsl-solver.js:16-40), one workgroup per tile in use
__device__ __forceinline__ void p2g_tile_body(const DParams* __restrict__ PP, const DGrid& G, uint32_t bid, uint32_t nbid) {
  const DParams& P = *PP;
  // per face: weight in the low 32 bits, velocity in the high 32 bits of one
  // 64-bit sum; weights are >= 0 and their sum < 2^31, so the halves never mix
  __shared__ unsigned long long sVW[3][WIN3];
  __shared__ int32_t sC[512];
  __shared__ uint32_t wslots[27];
  uint32_t vmaxBits = 0;
  uint32_t sc = G.ctl[4];
  for (uint32_t b = bid; b < sc; b += nbid) {
    uint32_t s = G.slotList[b], t = G.slotTile[s];
    int3 org;
    load_wslots(P, G, s, t, wslots, org);
    for (uint32_t q = threadIdx.x; q < WIN3; q += WG) {
      sVW[0][q] = 0ull; sVW[1][q] = 0ull; sVW[2][q] = 0ull;
    }
    for (uint32_t q = threadIdx.x; q < 512u; q += WG) sC[q] = 0;
    __syncthreads();
    uint32_t beg = G.binStart[t], end = G.binStart[t + 1];
    const bool wscaleSmall = P.scale <= 2147483520.0f;
    // With a power-of-two scale 2^0 .. 2^20, (w * vc) * scale equals w * (vc * scale)
    // bit for bit: scaling by 2^k commutes with rounding where w * vc is normal, and
    // where it is subnormal (or zero) both products round to the integer 0.
    const bool scalePow2 = (st_f2u(P.scale) & 0x807fffffu) == 0u && P.scale >= 1.0f && P.scale <= 1048576.0f;
    // the four waves take every fourth entry of each run of WG bin entries:
    // neighbouring entries tend to be neighbouring particles (last substep's
    // order), and lanes adding to the same LDS word in one instruction queue
    static_assert(WG == 256, "four waves of 64 cover a run of WG entries");
    for (uint32_t k0 = beg; k0 < end; k0 += WG) {
      const uint32_t k = k0 + (threadIdx.x & 63u) * 4u + (threadIdx.x >> 6);
      if (k >= end) continue;
      uint32_t i = G.binIdx[k];
      float4 p4 = G.pos[i];
      float4 v4 = G.vel[i];
      float3 p = make_float3(p4.x, p4.y, p4.z);
      float3 v = make_float3(v4.x, v4.y, v4.z);
      int3 cc = d_particleCell(P, p);
      vmaxBits = max(vmaxBits, st_f2u(f3len(v)));
      const WinAxes wa = win_axes(P, p, org);
#pragma unroll
      for (uint32_t axis = 0; axis < 3; axis++) {
        const WinAxis& ax = wa.x(axis);
        const WinAxis& ay = wa.y(axis);
        const WinAxis& az = wa.z(axis);
        float vc = f3get(v, axis);
        // every corner is added: a zero weight adds round(0) = 0 to an integer sum,
        // exactly as skipping it (wgsl-solver.js:33).  w is in [0, 1] (NaN for a
        // NaN position), so w * scale <= scale and the saturating clamp of
        // st_f2i_sat never acts when the scale is below it (wscaleSmall); likewise
        // |w * vc * scale| <= |vc * scale| below the clamp (fastV): the conversion
        // alone gives the same bits.  A NaN or infinite vc * scale takes the
        // written-out products.  On the fast path the scaled weight
        // wS = (fx * fy) * (fz * scale) serves both sums: with the power-of-two
        // scale it is w * scale bit for bit where w is normal, and wS * vc is
        // (w * vc) * scale; where w or w * vc is subnormal both forms round to 0.
        const float vcs = vc * P.scale;
        const bool fastV = scalePow2 && fabsf(vcs) < 2147483520.0f;
        auto corners = [&](auto fast) {
#pragma unroll
          for (uint32_t c = 0; c < 8; c++) {
            const uint32_t dx = c & 1u, dy = (c >> 1) & 1u, dz = (c >> 2) & 1u;
            int32_t av, aw;
            if constexpr (decltype(fast)::value) {
              const float wS = ((dx ? ax.f : ax.g) * (dy ? ay.f : ay.g)) * ((dz ? az.f : az.g) * P.scale);
              av = st_f2i_nosat(rintf(wS * vc));
              aw = st_f2i_nosat(rintf(wS));
            } else {
              float w = (dx ? ax.f : ax.g) * (dy ? ay.f : ay.g) * (dz ? az.f : az.g);
              av = st_f2i_sat(rintf(w * vc * P.scale));
              aw = wscaleSmall ? st_f2i_nosat(rintf(w * P.scale)) : st_f2i_sat(rintf(w * P.scale));
            }
            const uint32_t wi = win_idx(ax.w[dx], ay.w[dy], az.w[dz]);
            atomicAdd(&sVW[axis][wi], (unsigned long long)(uint32_t)aw + ((unsigned long long)(uint32_t)av << 32));
          }
        };
        if (fastV) corners(std::true_type{});
        else corners(std::false_type{});
      }
      atomicAdd(&sC[uint32_t(cc.x & 7) + 8u * (uint32_t(cc.y & 7) + 8u * uint32_t(cc.z & 7))], 1);
    }
    __syncthreads();
    // the window into HBM.  Counts and the faces of the tile's own cells are
    // stores (no other tile writes them: the neighbours' sums for the tile's
    // boundary cells wait in their halo records, which k_p2g_merge adds in);
    // the halo shell goes to this tile's record, except for cells of tiles
    // without storage (the spare slot), which every tile adds to directly.
    for (uint32_t q = threadIdx.x; q < 512u; q += WG) G.cnt[s * 512u + q] = sC[q];
    int32_t* rec = G.p2gHalo + size_t(s) * P2G_HALO_WORDS;
    for (uint32_t q = threadIdx.x; q < WIN3; q += WG) {
      const uint32_t wx = q % WIN, wy = (q / WIN) % WIN, wz = q / (WIN * WIN);
      const unsigned long long e[3] = {sVW[0][q], sVW[1][q], sVW[2][q]};
      const bool own = wx - 1u < 8u && wy - 1u < 8u && wz - 1u < 8u;
      if (own) {
        const uint32_t c = s * 512u + (wx - 1u) + 8u * ((wy - 1u) + 8u * (wz - 1u));
        for (int a = 0; a < 3; a++) {
          G.accV[c * 3 + a] = int32_t(uint32_t(e[a] >> 32));
          G.accW[c * 3 + a] = int32_t(uint32_t(e[a]));
        }
        continue;
      }
      const uint32_t ns = wslots[woff(int32_t(wx)) + 3 * (woff(int32_t(wy)) + 3 * woff(int32_t(wz)))];
      if (ns == WSLOT_OUT) continue;
      if (ns != 0u) {
        int32_t* r = rec + shell_idx(wx, wy, wz) * 6u;
        for (int a = 0; a < 3; a++) {
          r[2 * a] = int32_t(uint32_t(e[a] >> 32));
          r[2 * a + 1] = int32_t(uint32_t(e[a]));
        }
        continue;
      }
      const uint32_t c = wcell(wslots, int32_t(wx), int32_t(wy), int32_t(wz), org);
      for (int a = 0; a < 3; a++) {
        int32_t ww = int32_t(uint32_t(e[a])), vv = int32_t(uint32_t(e[a] >> 32));
        if (vv) atomicAdd(&G.accV[c * 3 + a], vv);
        if (ww) atomicAdd(&G.accW[c * 3 + a], ww);
      }
    }
    __syncthreads();
  }
  // fastest particle: one partial per workgroup (k_reduce_stats folds them into
  // ctl[6]); same-address atomics from every wave would serialise chip-wide
  __shared__ uint32_t wmax[WG / 64];
  vmaxBits = wave_max_u32(vmaxBits);
  if (lane_id() == 0) wmax[threadIdx.x >> 6] = vmaxBits;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t m = 0;
    for (int q = 0; q < WG / 64; q++) m = max(m, wmax[q]);
    G.part[bid] = m;
  }
}
__global__ void __launch_bounds__(WG) k_p2g_tile(const DParams* __restrict__ PP, DGrid G) {
  p2g_tile_body(PP, G, blockIdx.x, gridDim.x);
}

// After p2g: every tile in use adds to the faces of its boundary cells (a local
// coordinate at 0 or 7) the sums its neighbours' windows made for them, from
// the neighbours' halo records (tiles without storage ran no p2g).  A cell on
// the tile's faces, edges or corner lies in the windows of 1, 3 or 7
// neighbours.  Integer sums: any order gives the bits the atomics gave.
#define P2G_BOUNDARY 296  // cells of a tile with a local coordinate at 0 or 7
#define P2G_MERGE_WG 320   // one thread per boundary cell (five waves)
struct BoundaryCells {
  uint16_t l[P2G_BOUNDARY];
};
constexpr BoundaryCells boundary_cells() {
  BoundaryCells b{};
  int n = 0;
  for (int l = 0; l < 512; l++) {
    const int x = l & 7, y = (l >> 3) & 7, z = l >> 6;
    if (x == 0 || x == 7 || y == 0 || y == 7 || z == 0 || z == 7) b.l[n++] = uint16_t(l);
  }
  return b;
}
__constant__ BoundaryCells kBoundary = boundary_cells();

__device__ __forceinline__ void p2g_merge_body(const DGrid& G, uint32_t bid, uint32_t nbid) {
  const uint32_t sc = G.ctl[4];
  for (uint32_t b = bid; b < sc; b += nbid) {
    const uint32_t s = G.slotList[b];
    const uint32_t* nb = G.nbr + s * 27u;
    for (uint32_t q = threadIdx.x; q < P2G_BOUNDARY; q += blockDim.x) {
      const uint32_t l = kBoundary.l[q], x = l & 7u, y = (l >> 3) & 7u, z = l >> 6;
      const int32_t bx = x == 0u ? -1 : (x == 7u ? 1 : 0), by = y == 0u ? -1 : (y == 7u ? 1 : 0), bz = z == 0u ? -1 : (z == 7u ? 1 : 0);
      int32_t sum[6] = {0, 0, 0, 0, 0, 0};
      for (uint32_t k = 1; k < 8u; k++) {  // the neighbour offsets (k&1 ? bx : 0, k&2 ? by : 0, k&4 ? bz : 0)
        const int32_t dx = (k & 1u) ? bx : 0, dy = (k & 2u) ? by : 0, dz = (k & 4u) ? bz : 0;
        if ((k & 1u && !bx) || (k & 2u && !by) || (k & 4u && !bz)) continue;
        const uint32_t ns = nb[uint32_t(dx + 1) + 3u * uint32_t(dy + 1) + 9u * uint32_t(dz + 1)];
        if (ns == 0u || ns == WSLOT_OUT) continue;
        // this cell in the neighbour's window: local + 1 - 8 d per axis
        const uint32_t h = shell_idx(uint32_t(int32_t(x) + 1 - 8 * dx), uint32_t(int32_t(y) + 1 - 8 * dy), uint32_t(int32_t(z) + 1 - 8 * dz));
        const int32_t* r = G.p2gHalo + size_t(ns) * P2G_HALO_WORDS + h * 6u;
        for (int q = 0; q < 6; q++) sum[q] += r[q];
      }
      const uint32_t c = s * 512u + l;
      for (int a = 0; a < 3; a++) {
        if (sum[2 * a]) G.accV[c * 3 + a] += sum[2 * a];
        if (sum[2 * a + 1]) G.accW[c * 3 + a] += sum[2 * a + 1];
      }
    }
  }
}
__global__ void __launch_bounds__(P2G_MERGE_WG) k_p2g_merge(DGrid G) { p2g_merge_body(G, blockIdx.x, gridDim.x); }

// statistics partials -> control words: which 0: ctl[6] = max(ctl[6], vmax
// partials); which 1: ctl[0], ctl[1] = sums of the per-colour fluid partials
__device__ __forceinline__ void reduce_stats_body(const DGrid& G, uint32_t n, uint32_t which, uint32_t bid, uint32_t nbid) {
  __shared__ uint32_t ws[2][1024 / 64];
  uint32_t a = 0, b = 0;
  for (uint32_t i = threadIdx.x; i < n; i += 1024) {
    if (which == 0) a = max(a, G.part[i]);
    else { a += G.part[2 * i]; b += G.part[2 * i + 1]; }
  }
  if (which == 0) a = wave_max_u32(a);
  else { a = wave_sum_u32(a); b = wave_sum_u32(b); }
  if (lane_id() == 0) { ws[0][threadIdx.x >> 6] = a; ws[1][threadIdx.x >> 6] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t x = 0, y = 0;
    for (int q = 0; q < 1024 / 64; q++) {
      if (which == 0) x = max(x, ws[0][q]);
      else { x += ws[0][q]; y += ws[1][q]; }
    }
    if (which == 0) G.ctl[6] = max(G.ctl[6], x);
    else { G.ctl[0] = x; G.ctl[1] = y; }
  }
}
__global__ void __launch_bounds__(1024) k_reduce_stats(DGrid G, uint32_t n, uint32_t which) {
  reduce_stats_body(G, n, which, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// per active cell kernels: t -> slot slotList[t >> 9], local t & 511 (acell,
// wgsl-core.js:96-103)
#define FOR_ACTIVE(G)                                                                          \
  for (uint32_t t = bid * blockDim.x + threadIdx.x, tot_ = (G).ctl[4] * 512u; t < tot_; \
       t += nbid * blockDim.x)

// solidBuild (wgsl-solver.js:41-45)
__device__ __forceinline__ void solid_build_body(const DParams* __restrict__ PP, const DCols* __restrict__ C,
                                                 const DGrid& G, uint32_t bid, uint32_t nbid) {
  const DParams& P = *PP;
  FOR_ACTIVE(G) {
    uint32_t s = G.slotList[t >> 9], local = t & 511u, c = s * 512u + local;
    int3 p = d_slotCoord(P, G.slotTile, s, local);
    G.ctype[c] = (G.ctype[c] & 0xffff00ffu) | (uint32_t(d_colAt(P, C, G.sdf, d_cellCenter(P, p.x, p.y, p.z)) + 1) << 8);
  }
}
__global__ void k_solid_build(const DParams* __restrict__ PP, const DCols* __restrict__ C, DGrid G) {
  solid_build_body(PP, C, G, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ float3 d_velGVec(const DParams& P, const DGrid& G, float3 x);

// classify (wgsl-solver.js:46-74), plus the per-colour fluid counts of buildList
// (wgsl-solver.js:152-158) as statistics
__device__ __forceinline__ void classify_body(const DParams* __restrict__ PP, const DGrid& G, uint32_t bid, uint32_t nbid) {
  const DParams& P = *PP;
  uint32_t tot = G.ctl[4] * 512u;
  uint32_t stride = nbid * blockDim.x;
  uint32_t nred = 0, nblack = 0;
  for (uint32_t t = bid * blockDim.x + threadIdx.x; t < tot; t += stride) {
    uint32_t s = G.slotList[t >> 9], local = t & 511u, c = s * 512u + local;
    int3 a = d_slotCoord(P, G.slotTile, s, local);
    uint32_t ty = 0;
    uint32_t sb = G.ctype[c] & 0xffffff00u;
    int32_t n = G.cnt[c];
    if (!d_inDom(P, a)) ty = 0;
    else if ((sb & 0xff00u) != 0u) ty = 2;
    else if (P.ocean != 0u && d_layerW(P, d_cellCenter(P, a.x, a.y, a.z)) >= 0.5f) {
      // the ocean's boundary: a wall moving with the waves, under their surface
      float3 cx = d_cellCenter(P, a.x, a.y, a.z);
      ty = cx.y < d_oceanEta(P, cx.x, cx.z) ? 2u : 0u;
    } else if (P.nb != 0u && d_phiOfBits(sb) < -1.0f &&
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
  __shared__ uint32_t ws[2][WG / 64];
  nred = wave_sum_u32(nred);
  nblack = wave_sum_u32(nblack);
  if (lane_id() == 0) { ws[0][threadIdx.x >> 6] = nred; ws[1][threadIdx.x >> 6] = nblack; }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t r = 0, b = 0;
    for (int q = 0; q < WG / 64; q++) { r += ws[0][q]; b += ws[1][q]; }
    G.part[2 * bid] = r;
    G.part[2 * bid + 1] = b;
  }
}
__global__ void __launch_bounds__(WG) k_classify(const DParams* __restrict__ PP, DGrid G) {
  classify_body(PP, G, blockIdx.x, gridDim.x);
}

// normBody / normFaces (wgsl-solver.js:75-105)
__device__ __forceinline__ void norm_faces_body(const DParams* __restrict__ PP, const DCols* __restrict__ C, const DGrid& G,
                                                uint32_t bid, uint32_t nbid) {
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
        else if (P.ocean != 0u && ((tl == 2u && d_inDom(P, lo)) || (th == 2u && d_inDom(P, at))))
          sv = f3get(d_oceanVel(P, fp), axis);  // the ocean's moving wall
        v = sv;
        old = sv;
      }
      G.accW[f] = (int32_t)st_f2u(old);
      G.accV[f] = (int32_t)st_f2u(v);
    }
  }
}
__global__ void k_norm_faces(const DParams* __restrict__ PP, const DCols* __restrict__ C, DGrid G) {
  norm_faces_body(PP, C, G, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ float gvR(const DGrid& G, uint32_t f) { return st_u2f((uint32_t)G.accV[f]); }
__device__ __forceinline__ float gv0R(const DGrid& G, uint32_t f) { return st_u2f((uint32_t)G.accW[f]); }

// SOR stencil byte: bit 7 fluid, bits 0-2 the count of non-solid neighbours
#define STEN_FLUID 0x80u

// divergence (wgsl-solver.js:106-123); also records the SOR stencil of the cell,
// which depends only on the cell types fixed by classify.
__device__ __forceinline__ void divergence_body(const DParams* __restrict__ PP, const DGrid& G, uint32_t bid, uint32_t nbid) {
  const DParams& P = *PP;
  FOR_ACTIVE(G) {
    uint32_t s = G.slotList[t >> 9], local = t & 511u, c = s * 512u + local;
    uint32_t ct = G.ctype[c];
    if ((ct & 0xffu) != 1u) {
      G.rhs[c] = 0.0f;
      G.sten[c] = 0;
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
    uint32_t nns = 0;
    for (int q = 0; q < 6; q++) {
      if (tys[q] == 0u) interior = false;
      if (tys[q] != 2u) nns++;
    }
    float dens = float(G.cnt[c]) / P.rest - 1.0f;
    float tgt = P.comp * (interior ? dens : fmaxf(dens, 0.0f));
    if (P.nb != 0u && d_phiOfBits(ct) < -1.0f) tgt = 0.0f;
    G.rhs[c] = (tgt - div) * P.h * P.h / P.dt;
    G.sten[c] = uint8_t(STEN_FLUID | nns);
  }
}
__global__ void k_divergence(const DParams* __restrict__ PP, DGrid G) {
  divergence_body(PP, G, blockIdx.x, gridDim.x);
}

// relax / rbRed / rbBlack (wgsl-solver.js:124-140, 161-164; loop at
// sim.js:291-293): one red-black iteration per launch, by tiles.  A workgroup
// stages the pressure of a 12^3 window (the tile and two cells of halo) and the
// rhs and stencil of the 10^3 window in LDS, relaxes the red cells of the tile
// and of the one-cell slabs across its six faces, then the black cells of the
// tile, and writes the tile to the other pressure buffer.  A red cell reads only black neighbours and a black cell only
// red ones, so the black sweep sees exactly the red values the neighbouring
// tiles compute, and the result is the reference's red-black iterate bit for bit.
//
// Every non-fluid cell of the window holds +0.0 (classify zeroes the pressure
// of every non-fluid cell, and missing tiles load as zero), so all six
// neighbours are summed in the reference's order (x-, x+, y-, y+, z-, z+): a
// non-fluid neighbour adds +0.0, the reference's partial sums start at +0.0 and
// are never -0.0, and x + 0 = x for every other x.  The quotient by the count
// of non-solid neighbours is st_div_small, equal to the division.
template <int W>
__device__ __forceinline__ void relax_lds(float* sp, int32_t p, uint32_t st, float rhs, float omega, const float* ytab) {
  // st: the stencil byte (bit 7 fluid, bits 0-2 non-solid neighbours)
  uint32_t nns = st & 7u;
  if (nns == 0u) {
    sp[p] = 0.0f;
    return;
  }
  float n0 = sp[p - 1], n1 = sp[p + 1], n2 = sp[p - W], n3 = sp[p + W], n4 = sp[p - W * W], n5 = sp[p + W * W];
  float y = ytab[nns], pc = sp[p];
  float sum = 0.0f;
  sum = sum + n0;
  sum = sum + n1;
  sum = sum + n2;
  sum = sum + n3;
  sum = sum + n4;
  sum = sum + n5;
  float pg = st_div_small(rhs + sum, float(nns), y);
  sp[p] = pc + omega * (pg - pc);
}

// a window slot that holds no data: outside the box, or a tile without storage
// (the spare slot 0 stands in for it)
__device__ __forceinline__ bool sor_missing(uint32_t slot) { return slot + 2u <= 2u; }

// The launch is persistent: about eight workgroups per CU, each walking tiles
// b, b + G, b + 2G, ... of its share of slotList.  While a workgroup relaxes
// tile b in LDS, the window rows of tile b + G are already in
// flight into registers (thread r < 144 holds pressure row r of the 12^3
// window; thread 144 + r < 244 holds rhs and stencil row r of the 10^3
// window), and the 27 slots of tile b + 2G are being fetched.  Tiles are
// independent within the launch (they read pin, write pout), so the pipeline
// only reorders loads.
// A thread holds either a pressure row (threads 0..143) or an rhs and stencil
// row (threads 144..243) in the same 13 registers:
//   pressure: q0, q1 the 8 cells of the row's own tile, q2 = (two cells to the
//             left, two to the right);
//   rhs/sten: q0, q1 rhs of the own tile, q2 = (rhs left, rhs right, stencil
//             bytes of the own tile as two words), e = stencil left | right << 8.
struct SorRows {
  uint4 q0, q1, q2;
  uint32_t e;
};

__device__ __forceinline__ SorRows sor_fetch(const DGrid& G, const float* __restrict__ pin, const uint32_t* ws, uint32_t tid) {
  const uint4 z4 = make_uint4(0u, 0u, 0u, 0u);
  SorRows R{z4, z4, z4, 0u};
  if (tid < 144u) {
    int32_t ly = int32_t(tid % 12u) - 2, lz = int32_t(tid / 12u) - 2;
    uint32_t rowoff = 8u * (uint32_t(ly & 7) + 8u * uint32_t(lz & 7));
    int32_t so = 3 * (((ly >> 3) + 1) + 3 * ((lz >> 3) + 1));
    uint32_t sl = ws[so], sc = ws[so + 1], sh = ws[so + 2];
    const uint32_t* src = reinterpret_cast<const uint32_t*>(pin);
    uint2 l = sor_missing(sl) ? make_uint2(0u, 0u) : *(const uint2*)(src + sl * 512u + rowoff + 6u);
    uint2 h = sor_missing(sh) ? make_uint2(0u, 0u) : *(const uint2*)(src + sh * 512u + rowoff);
    if (!sor_missing(sc)) {
      R.q0 = *(const uint4*)(src + sc * 512u + rowoff);
      R.q1 = *(const uint4*)(src + sc * 512u + rowoff + 4u);
    }
    R.q2 = make_uint4(l.x, l.y, h.x, h.y);
  } else if (tid < 244u) {
    uint32_t r = tid - 144u;
    int32_t ly = int32_t(r % 10u) - 1, lz = int32_t(r / 10u) - 1;
    uint32_t rowoff = 8u * (uint32_t(ly & 7) + 8u * uint32_t(lz & 7));
    int32_t so = 3 * (((ly >> 3) + 1) + 3 * ((lz >> 3) + 1));
    uint32_t sl = ws[so], sc = ws[so + 1], sh = ws[so + 2];
    const uint32_t* rhs = reinterpret_cast<const uint32_t*>(G.rhs);
    uint32_t rl = 0u, rh = 0u, tl = 0u, th = 0u;
    if (!sor_missing(sl)) {
      rl = rhs[sl * 512u + rowoff + 7u];
      tl = G.sten[sl * 512u + rowoff + 7u];
    }
    if (!sor_missing(sh)) {
      rh = rhs[sh * 512u + rowoff];
      th = G.sten[sh * 512u + rowoff];
    }
    uint2 t = make_uint2(0u, 0u);
    if (!sor_missing(sc)) {
      R.q0 = *(const uint4*)(rhs + sc * 512u + rowoff);
      R.q1 = *(const uint4*)(rhs + sc * 512u + rowoff + 4u);
      t = *(const uint2*)(G.sten + sc * 512u + rowoff);
    }
    R.q2 = make_uint4(rl, rh, t.x, t.y);
    R.e = tl | (th << 8);
  }
  return R;
}
__device__ __forceinline__ void sor_row8(float* d, uint4 a, uint4 b) {
  d[0] = __uint_as_float(a.x); d[1] = __uint_as_float(a.y); d[2] = __uint_as_float(a.z); d[3] = __uint_as_float(a.w);
  d[4] = __uint_as_float(b.x); d[5] = __uint_as_float(b.y); d[6] = __uint_as_float(b.z); d[7] = __uint_as_float(b.w);
}
// LDS image: sp the 12^3 pressure window; srs the 10^3 window of (rhs bits, stencil byte)
__device__ __forceinline__ void sor_store(float* sp, uint2* srs, uint32_t tid, SorRows R) {
  if (tid < 144u) {
    float* d = sp + 12u * tid;
    d[0] = __uint_as_float(R.q2.x);
    d[1] = __uint_as_float(R.q2.y);
    sor_row8(d + 2, R.q0, R.q1);
    d[10] = __uint_as_float(R.q2.z);
    d[11] = __uint_as_float(R.q2.w);
  } else if (tid < 244u) {
    uint2* d = srs + 10u * (tid - 144u);
    d[0] = make_uint2(R.q2.x, R.e & 0xffu);
    d[1] = make_uint2(R.q0.x, R.q2.z & 0xffu);
    d[2] = make_uint2(R.q0.y, (R.q2.z >> 8) & 0xffu);
    d[3] = make_uint2(R.q0.z, (R.q2.z >> 16) & 0xffu);
    d[4] = make_uint2(R.q0.w, R.q2.z >> 24);
    d[5] = make_uint2(R.q1.x, R.q2.w & 0xffu);
    d[6] = make_uint2(R.q1.y, (R.q2.w >> 8) & 0xffu);
    d[7] = make_uint2(R.q1.z, (R.q2.w >> 16) & 0xffu);
    d[8] = make_uint2(R.q1.w, R.q2.w >> 24);
    d[9] = make_uint2(R.q2.y, R.e >> 8);
  }
}

// LDS of one tile pipeline (256 threads)
struct SorPipeLds {
  float sp[SW3_1];
  uint2 srs[WIN3];
  uint32_t wsl[2][27];
  uint32_t redIdx[SOR_NRED];  // red cells relaxed per tile: 12^3 index | 10^3 index << 16
  float ytab[8];              // 1 / n, correctly rounded
};

// The tables of the pipeline (tid: 0..255 within its group of 256 threads).
__device__ __forceinline__ void sor_pipe_init(SorPipeLds& L, uint32_t tid) {
  // The black sweep reads red cells of the tile and of the six one-cell slabs
  // across its faces (never edges or corners): 256 + 6 x 32 red cells.  Tile
  // coordinates (x, y, z) in [-1, 8]; red means x + y + z even (the tile origin
  // is even).
  for (uint32_t m = tid; m < SOR_NRED; m += WG) {
    int32_t x, y, z;
    if (m < 256u) {
      y = int32_t((m >> 2) & 7u); z = int32_t(m >> 5);
      x = 2 * int32_t(m & 3u) + ((y + z) & 1);
    } else {
      uint32_t f = (m - 256u) >> 5, j = (m - 256u) & 31u;
      int32_t e = (f & 1u) ? 8 : -1, u = int32_t(j >> 2), v = 2 * int32_t(j & 3u);
      if (f < 2u) { x = e; z = u; y = v + ((e + u) & 1); }
      else if (f < 4u) { y = e; z = u; x = v + ((e + u) & 1); }
      else { z = e; y = u; x = v + ((e + u) & 1); }
    }
    L.redIdx[m] = uint32_t((x + 2) + 12 * ((y + 2) + 12 * (z + 2))) | (uint32_t((x + 1) + 10 * ((y + 1) + 10 * (z + 1))) << 16);
  }
  if (tid < 8u) L.ytab[tid] = (tid >= 1u && tid <= 6u) ? 1.0f / float(tid) : 0.0f;
}

// One red-black iteration over the tiles of virtual workgroup `vid` of `vgrid`
// (a multiple of 8; vid runs on XCD vid % 8).  Each XCD walks one contiguous
// eighth of the tile-ordered active list, so x and y neighbours of a tile are
// processed by the same XCD at about the same time and their halo rows hit its
// L2.  The loop runs `kmax` rounds for every group of the workgroup (so all
// groups meet at each __syncthreads); a group past its share idles.
__device__ __forceinline__ void sor_pipe_pass(const DGrid& G, float omega, const float* __restrict__ pin, float* __restrict__ pout,
                                              uint32_t vid, uint32_t vgrid, uint32_t tid, SorPipeLds& L) {
  const uint32_t nact = G.ctl[4], gs = vgrid >> 3, xcd = vid & 7u;
  const uint32_t per = (nact + 7u) >> 3;
  const uint32_t sc = min(nact, (xcd + 1u) * per);
  const uint32_t kmax = (per + gs - 1u) / gs;
  uint32_t b = xcd * per + (vid >> 3);
  // prologue: slots of tiles b and b + gs, window rows of tile b
  if (tid < 27u) {
    L.wsl[0][tid] = b < sc ? G.nbrL[b * 27u + tid] : 0u;
    L.wsl[1][tid] = b + gs < sc ? G.nbrL[(b + gs) * 27u + tid] : 0u;
  }
  __syncthreads();
  SorRows R{};
  if (b < sc) R = sor_fetch(G, pin, L.wsl[0], tid);
  uint32_t cur = 0;
  for (uint32_t k = 0; k < kmax; k++, b += gs, cur ^= 1u) {
    const bool act = b < sc, more = b + gs < sc;
    const uint32_t s = L.wsl[cur][13];
    uint32_t ahead = 0u;
    if (tid < 27u && b + 2u * gs < sc) ahead = G.nbrL[(b + 2u * gs) * 27u + tid];
    if (act) sor_store(L.sp, L.srs, tid, R);
    __syncthreads();
    if (more) R = sor_fetch(G, pin, L.wsl[cur ^ 1u], tid);
    if (act) {  // red cells the black sweep reads
      for (uint32_t m = tid; m < SOR_NRED; m += WG) {
        uint32_t e = L.redIdx[m];
        uint2 rs = L.srs[e >> 16];
        if (rs.y & STEN_FLUID) relax_lds<12>(L.sp, int32_t(e & 0xffffu), rs.y, __uint_as_float(rs.x), omega, L.ytab);
      }
    }
    __syncthreads();
    if (act) {  // black cells of the tile, one per thread
      int32_t y = int32_t((tid >> 2) & 7u), z = int32_t(tid >> 5);
      int32_t x = 2 * int32_t(tid & 3u) + ((y + z + 1) & 1);
      uint2 rs = L.srs[(x + 1) + 10 * ((y + 1) + 10 * (z + 1))];
      if (rs.y & STEN_FLUID) relax_lds<12>(L.sp, (x + 2) + 12 * ((y + 2) + 12 * (z + 2)), rs.y, __uint_as_float(rs.x), omega, L.ytab);
    }
    __syncthreads();
    if (act) {
      for (uint32_t q = tid; q < 512u; q += WG) {
        uint32_t x = q & 7u, y = (q >> 3) & 7u, z = q >> 6;
        pout[s * 512u + q] = L.sp[(x + 2) + 12 * ((y + 2) + 12 * (z + 2))];
      }
    }
    // slots of tile b + 2gs go where tile b's were (read above into s)
    if (tid < 27u) L.wsl[cur][tid] = ahead;
    __syncthreads();
  }
}

// One red-black iteration per launch (the grid a multiple of 8 workgroups, as
// many as are resident), skipped when the resident solve has done this stage.
__device__ __forceinline__ void sor_pipe_body(const DParams* __restrict__ PP, const DGrid& G, const float* __restrict__ pin,
                                              float* __restrict__ pout, uint32_t seq, uint32_t bid, uint32_t nbid) {
  __shared__ SorPipeLds L;
  if (G.sorBar[COOP_MAX_WG] == seq) return;
  sor_pipe_init(L, threadIdx.x);
  sor_pipe_pass(G, PP->omega, pin, pout, bid, nbid, threadIdx.x, L);
}
__global__ void __launch_bounds__(WG) k_sor_pipe(const DParams* __restrict__ PP, DGrid G, const float* __restrict__ pin,
                                                 float* __restrict__ pout, uint32_t seq) {
  sor_pipe_body(PP, G, pin, pout, seq, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// Resident SOR: every iteration of one pressure solve in one cooperative launch,
// one workgroup of COOP_WG threads per CU.  Workgroup w owns the w-th contiguous
// range of the active tiles in Morton order (compact blobs of about nact / grid
// tiles) and keeps, for the whole solve, their pressure in LDS, split by colour
// (sP), and the rhs and stencil of the cells each thread relaxes in registers.
// Only the face layers cross workgroups: after relaxing colour c a workgroup
// writes the colour-c cells of every face whose neighbour tile another
// workgroup owns to sorHalo[c], each tagged with the half-sweep, and before
// relaxing the other colour it waits for and copies the colour-c cells across
// the faces toward other workgroups' tiles into sH (faces toward its own tiles
// are read in place).  No grid barrier: a workgroup overwrites its colour-c
// entries two half-sweeps later, only after it has read the next half-sweep's
// entries of every workgroup that reads it (tiles that touch read each other),
// which those write after their reads are done.  The arithmetic is
// relax_lds's, cell for cell, so the iterate is the reference's bit for bit.  When the active tiles exceed COOP_TMAX per
// workgroup the same launch streams tile windows instead (below).

// index of tile-local cell (x, y, z) in the colour-split layout: colour
// (x + y + z) & 1 (the tile origin is even, so this is the global colour), then
// (z, y, x / 2)
__device__ __forceinline__ uint32_t coop_cidx(int32_t x, int32_t y, int32_t z) {
  return (uint32_t(x + y + z) & 1u) * 256u + uint32_t(z * 8 + y) * 4u + uint32_t(x >> 1);
}
// faces 0..5 = x-, x+, y-, y+, z-, z+; the nbr column of the tile across face f
__device__ __forceinline__ uint32_t coop_faceq(uint32_t f) {
  return f == 0u ? 12u : f == 1u ? 14u : f == 2u ? 10u : f == 3u ? 16u : f == 4u ? 4u : 22u;
}
// cell e (0..31) of colour c on tile-local face f: (u, v) are the in-plane
// coordinates in axis order, e = 4 v + u / 2, and the normal coordinate is 0 on
// a low face and 7 on a high one
__device__ __forceinline__ int3 coop_face_cell(uint32_t f, uint32_t e, uint32_t c) {
  int32_t n0 = (f & 1u) ? 7 : 0, v = int32_t(e >> 2);
  int32_t u = 2 * int32_t(e & 3u) + ((int32_t(c) + n0 + v) & 1);
  uint32_t a = f >> 1;
  return a == 0u ? make_int3(n0, u, v) : (a == 1u ? make_int3(u, n0, v) : make_int3(u, v, n0));
}

#define COOP_NB_NONE 0xffffffffu
#define COOP_NB_FOREIGN 0x80000000u
#define COOP_SPIN_LIMIT GB_SPIN_LIMIT

// Barrier over the resident SOR's ng workgroups of one sim (grid_barrier.h):
// flags in sorBar[0 .. ng), time-outs counted in sorBar[COOP_MAX_WG + 1], the
// release word sorBar[COOP_MAX_WG + 2].
__device__ __forceinline__ bool coop_barrier(const DGrid& G, uint32_t epoch, uint32_t* sFail, uint32_t w, uint32_t ng) {
  return gb_barrier(G.sorBar, G.sorBar + COOP_MAX_WG + 2, G.sorBar + COOP_MAX_WG + 1, epoch, sFail, w, ng);
}

// sH holds, per resident tile and face, the 32 cells across that face that the
// face relaxation reads from another workgroup's tile (or zeros, set once by
// coop_zero_none, where there is no tile; faces toward our own tiles are read
// from those tiles directly).  coop_fill_start reads the foreign faces from the
// pressure the solve starts from; coop_fill_foreign takes them from the halo
// entries of the half-sweep tagged `tag`: each entry is one 64-bit word, the f32
// value in the low half and the tag of the half-sweep that wrote it in the high
// half, stored and loaded as a whole, so an entry with the right tag holds that
// half-sweep's value and no flag or fence is needed.  A thread keeps polling its
// entries (COOP_FOR_PER_THREAD in flight together) until their tags match, at
// most COOP_SPIN_LIMIT rounds; a time-out sets *sFail.
#define COOP_FOR_PER_THREAD 4
__device__ __forceinline__ void coop_zero_none(float* sH, const uint32_t* sNb, uint32_t T) {
  for (uint32_t q = threadIdx.x; q < T * 192u; q += COOP_WG)
    if (sNb[q >> 5] == COOP_NB_NONE) sH[q] = 0.0f;
}
__device__ __forceinline__ void coop_fill_start(const float* __restrict__ pin, float* sH, const uint32_t* sNb, const uint32_t* sFor,
                                                uint32_t nFor, uint32_t c) {
  for (uint32_t q = threadIdx.x; q < nFor * 32u; q += COOP_WG) {
    uint32_t hf = sFor[q >> 5], e = q & 31u, f = hf % 6u;
    uint32_t ns = sNb[hf] & ~COOP_NB_FOREIGN;
    int3 a = coop_face_cell(f ^ 1u, e, c);
    sH[hf * 32u + e] = pin[ns * 512u + uint32_t(a.x + 8 * (a.y + 8 * a.z))];
  }
}
__device__ __forceinline__ size_t coop_halo_idx(uint32_t c, uint32_t scap, uint32_t slot, uint32_t f, uint32_t e) {
  return ((size_t(c) * scap + slot) * 6u + f) * 32u + e;
}
__device__ __forceinline__ void coop_halo_put(const DGrid& G, size_t idx, float v, uint32_t tag) {
  const uint64_t w = (uint64_t(tag) << 32) | uint64_t(__float_as_uint(v));
  __hip_atomic_store(&G.sorHalo[idx], w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// k-th foreign face: its 32 entries are polled at sGetG[k] + cOff + e (cOff
// the colour's offset) and copied to sH[sFor[k] * 32 + e]
__device__ __forceinline__ void coop_fill_foreign(const DGrid& G, float* sH, const uint32_t* sFor, const uint32_t* sGetG,
                                                  uint32_t nFor, uint32_t cOff, uint32_t tag, uint32_t* sFail) {
  for (uint32_t q0 = 0; q0 < nFor * 32u; q0 += COOP_FOR_PER_THREAD * COOP_WG) {
    uint64_t v[COOP_FOR_PER_THREAD];
    uint32_t at[COOP_FOR_PER_THREAD];
#pragma unroll
    for (uint32_t i = 0; i < COOP_FOR_PER_THREAD; i++) {
      uint32_t q = q0 + threadIdx.x + i * COOP_WG;
      at[i] = 0;
      v[i] = uint64_t(tag) << 32;
      if (q < nFor * 32u) {
        at[i] = sGetG[q >> 5] + cOff + (q & 31u);
        v[i] = __hip_atomic_load(&G.sorHalo[at[i]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    for (uint32_t spins = 0;; spins++) {
      bool ready = true;
#pragma unroll
      for (uint32_t i = 0; i < COOP_FOR_PER_THREAD; i++) ready = ready && uint32_t(v[i] >> 32) == tag;
      if (ready) break;
      if (spins >= COOP_SPIN_LIMIT) {
        *sFail = 1u;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
#pragma unroll
      for (uint32_t i = 0; i < COOP_FOR_PER_THREAD; i++)
        if (uint32_t(v[i] >> 32) != tag) v[i] = __hip_atomic_load(&G.sorHalo[at[i]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (uint32_t i = 0; i < COOP_FOR_PER_THREAD; i++) {
      uint32_t q = q0 + threadIdx.x + i * COOP_WG;
      if (q < nFor * 32u) sH[sFor[q >> 5] * 32u + (q & 31u)] = __uint_as_float(uint32_t(v[i]));
    }
  }
}

// The resident solve relaxes rows: item lt * 64 + r is row r = 8 z + y of
// resident tile lt, whose colour-c cells x = x0, x0 + 2, x0 + 4, x0 + 6
// (x0 = (y + z + c) & 1) lie side by side in the colour-split layout, one
// float4.  Their neighbours across y and z are the same four positions of the
// other colour's rows r -+ 1 and r -+ 8 (the tile's origin is even, so the
// other colour's cells in those rows sit at the same x), or of the tile or
// halo copy across that face; across x they are the other colour's row r
// itself shifted by one place, with one cell from across an x face.
#define COOP_J ((COOP_TMAX * 64 + COOP_WG - 1) / COOP_WG)  // row items per thread and colour
static_assert(COOP_TMAX * 64 <= 4 * COOP_WG, "the item scan gives each thread four items");

// one cell of relax_lds: the same sum, quotient and over-relaxation, bit for
// bit; a cell outside the fluid keeps its value, a fluid cell with no open
// neighbour becomes 0.  The quotient's exact fallback runs only for the cells
// that need it (st_div_small's range test), not for cells that keep a value.
__device__ __forceinline__ float coop_cell(float pc, float rhs, uint32_t st, float n0, float n1, float n2, float n3, float n4,
                                           float n5, float omega, const float* ytab) {
  const uint32_t nns = st & 7u;
  float sum = 0.0f;
  sum = sum + n0;
  sum = sum + n1;
  sum = sum + n2;
  sum = sum + n3;
  sum = sum + n4;
  sum = sum + n5;
  const float a = rhs + sum, n = float(nns), y = ytab[nns];
  const float q = a * y;
  const float rr = fmaf(-q, n, a);
  float pg = fmaf(rr, y, q);
  const uint32_t ua = st_f2u(a) & 0x7fffffffu;
  if (nns != 0u && !(ua - 0x01800000u < 0x7e000000u)) pg = a / n;
  float out = pc + omega * (pg - pc);
  if (nns == 0u) out = (st & STEN_FLUID) ? 0.0f : pc;
  return out;
}

// one row item (colour c): rhs and stencil bytes of its four cells in rhs4 / st4
// (byte k for cell k).  Reads only the other colour and writes only this row.
__device__ __forceinline__ void coop_relax_row(float* sP, const float* sH, const uint32_t* sNb, uint32_t item, uint32_t c,
                                               float4 rhs4, uint32_t st4, float omega, const float* ytab) {
  const uint32_t lt = item >> 6, r = item & 63u, y = r & 7u, z = r >> 3, x0 = (y + z + c) & 1u;
  const uint32_t oc = (c ^ 1u) * 256u;
  // offsets as 24-bit products (lt < 64): no 32- or 64-bit multiplies
  float* cen = sP + ((lt << 9) + c * 256u + r * 4u);
  const float* ob = sP + ((lt << 9) + oc);
  const float* hb = sH + __umul24(lt, 192u);
  const uint32_t* nbs = sNb + __umul24(lt, 6u);
  // the address of the neighbour across face f: index w in the tile there when
  // this workgroup holds it, else index h in the face's halo copy
  auto across = [&](uint32_t f, uint32_t w, uint32_t h) -> const float* {
    const uint32_t nb = nbs[f];
    return (nb & COOP_NB_FOREIGN) ? hb + (f * 32u + h) : sP + ((nb << 9) + oc + w);
  };
  const float4 O = *reinterpret_cast<const float4*>(ob + r * 4u);
  const float A = *across(x0, r * 4u + (x0 ? 0u : 3u), z * 4u + (y >> 1));
  const float4 Ym = *reinterpret_cast<const float4*>(y > 0u ? ob + (r - 1u) * 4u : across(2u, (z * 8u + 7u) * 4u, z * 4u));
  const float4 Yp = *reinterpret_cast<const float4*>(y < 7u ? ob + (r + 1u) * 4u : across(3u, z * 32u, z * 4u));
  const float4 Zm = *reinterpret_cast<const float4*>(z > 0u ? ob + (r - 8u) * 4u : across(4u, (56u + y) * 4u, y * 4u));
  const float4 Zp = *reinterpret_cast<const float4*>(z < 7u ? ob + (r + 8u) * 4u : across(5u, y * 4u, y * 4u));
  const float4 pc = *reinterpret_cast<const float4*>(cen);
  // x neighbours of cell k: x0 = 0 has (k ? O[k-1] : A, O[k]); x0 = 1 has (O[k], k < 3 ? O[k+1] : A)
  const float m0 = x0 ? O.x : A, m1 = x0 ? O.y : O.x, m2 = x0 ? O.z : O.y, m3 = x0 ? O.w : O.z;
  const float p0 = x0 ? O.y : O.x, p1 = x0 ? O.z : O.y, p2 = x0 ? O.w : O.z, p3 = x0 ? A : O.w;
  float4 o;
  o.x = coop_cell(pc.x, rhs4.x, st4 & 0xffu, m0, p0, Ym.x, Yp.x, Zm.x, Zp.x, omega, ytab);
  o.y = coop_cell(pc.y, rhs4.y, (st4 >> 8) & 0xffu, m1, p1, Ym.y, Yp.y, Zm.y, Zp.y, omega, ytab);
  o.z = coop_cell(pc.z, rhs4.z, (st4 >> 16) & 0xffu, m2, p2, Ym.z, Yp.z, Zm.z, Zp.z, omega, ytab);
  o.w = coop_cell(pc.w, rhs4.w, st4 >> 24, m3, p3, Ym.w, Yp.w, Zm.w, Zp.w, omega, ytab);
  *reinterpret_cast<float4*>(cen) = o;
}

// whether row item (lt, r) of colour c has a cell on a face whose neighbour
// tile another workgroup holds (those cells are sent after each half-sweep)
__device__ __forceinline__ bool coop_row_early(const uint32_t* sNb, uint32_t lt, uint32_t r, uint32_t c) {
  const uint32_t y = r & 7u, z = r >> 3, x0 = (y + z + c) & 1u;
  auto foreign = [&](uint32_t f) {
    const uint32_t nb = sNb[lt * 6u + f];
    return nb != COOP_NB_NONE && (nb & COOP_NB_FOREIGN) != 0u;
  };
  return foreign(x0) || (y == 0u && foreign(2u)) || (y == 7u && foreign(3u)) || (z == 0u && foreign(4u)) ||
         (z == 7u && foreign(5u));
}

// LDS of the resident mode
struct CoopLds {
  float sP[COOP_TMAX * 512];
  float sH[COOP_TMAX * 192];
  uint32_t sNb[COOP_TMAX * 6];
  uint32_t sSlot[COOP_TMAX];
  uint32_t sFor[COOP_TMAX * 6];  // (tile, face) pairs whose neighbour another workgroup owns
  // per foreign face k (sFor order): lt * 512 | f << 16 of our face, the halo index of
  // our face's entry 0 (colour 0), and of the entry 0 of the neighbour's face we read
  uint32_t sPutL[COOP_TMAX * 6];
  uint32_t sPutG[COOP_TMAX * 6];
  uint32_t sGetG[COOP_TMAX * 6];
  uint16_t sFC[6 * 2 * 32];  // colour-split index of cell e of colour c on face f, at (2 f + c) * 32 + e
};
static_assert(sizeof(float) * COOP_TMAX * 192 >= 2 * COOP_TMAX * 64 * sizeof(uint16_t), "the row lists fit in the halo copy");
constexpr size_t kCoopLdsBytes = sizeof(CoopLds) > 4 * sizeof(SorPipeLds) ? sizeof(CoopLds) : 4 * sizeof(SorPipeLds);
static_assert(kCoopLdsBytes + 256 <= 160 * 1024, "the resident solve's LDS fits a CU");

// The resident solve of one sim by its workgroups w = 0 .. ng-1 (the whole
// launch for one sim, or its range of a batched launch).
__device__ __forceinline__ void sor_resident_body(const DParams* __restrict__ PP, const DGrid& G, uint32_t iters, uint32_t seq,
                                                  uint32_t stream, uint32_t w, uint32_t ng) {
  __shared__ __align__(16) unsigned char smem[kCoopLdsBytes];
  __shared__ uint32_t nFor;
  __shared__ float ytab[8];
  __shared__ uint32_t sFail;
  const uint32_t nact = G.ctl[4], tid = threadIdx.x;
  if (nact == 0u || ng > COOP_MAX_WG) return;
  const float omega = PP->omega;
  const uint32_t tag0 = PP->sorTag;
  if (tid == 0u) sFail = 0u;
  if (stream || nact > ng * COOP_TMAX) {
    // Streaming mode, for more tiles than fit in LDS: four tile pipelines of 256
    // threads per workgroup run the per-iteration pass (as k_sor_pipe, virtual
    // workgroup g * grid + w), with a grid barrier between iterations.
    const uint32_t g = tid >> 8, gt = tid & 255u;
    SorPipeLds& L = reinterpret_cast<SorPipeLds*>(smem)[g];
    sor_pipe_init(L, gt);
    uint32_t epoch = 0;
    for (uint32_t it = 0; it < iters; it++) {
      const float* pin = (it & 1u) ? G.pres2 : G.pres;
      float* pout = (it & 1u) ? G.pres : G.pres2;
      sor_pipe_pass(G, omega, pin, pout, g * ng + w, 4u * ng, gt, L);
      if (it + 1u < iters && !coop_barrier(G, ++epoch, &sFail, w, ng)) return;
    }
    if (iters & 1u) {  // the result sits in pres2: copy back the tiles this pipeline wrote
      const uint32_t vid = g * ng + w, gs = (4u * ng) >> 3, xcd = vid & 7u, per = (nact + 7u) >> 3;
      const uint32_t sc = min(nact, (xcd + 1u) * per);
      for (uint32_t b = xcd * per + (vid >> 3); b < sc; b += gs) {
        uint32_t sl = G.slotList[b];
        for (uint32_t q = gt; q < 512u; q += WG) G.pres[sl * 512u + q] = G.pres2[sl * 512u + q];
      }
    }
    if (w == 0u && tid == 0u) {
      G.sorBar[COOP_MAX_WG] = seq;
      G.sorBar[COOP_MAX_WG + 3] = 2u;
    }
    return;
  }
  CoopLds& CL = *reinterpret_cast<CoopLds*>(smem);
  float* sP = CL.sP;
  float* sH = CL.sH;
  uint32_t* sNb = CL.sNb;
  uint32_t* sSlot = CL.sSlot;
  uint32_t* sFor = CL.sFor;
  const uint32_t scap = PP->scap, cStride = scap * 192u;  // halo entries of one colour
  const uint32_t beg = uint32_t(uint64_t(nact) * w / ng), T = uint32_t(uint64_t(nact) * (w + 1u) / ng) - beg;
  float* pres = G.pres;
  if (tid == 0u) nFor = 0u;
  if (tid < 8u) ytab[tid] = (tid >= 1u && tid <= 6u) ? 1.0f / float(tid) : 0.0f;
  if (tid < T) sSlot[tid] = G.coopList[beg + tid];
  __syncthreads();
  // the tile across each face: one of ours (its local index), another
  // workgroup's (its slot), or none
  for (uint32_t q = tid; q < T * 6u; q += COOP_WG) {
    uint32_t lt = q / 6u, f = q - lt * 6u;
    uint32_t ns = G.nbr[sSlot[lt] * 27u + coop_faceq(f)], code = COOP_NB_NONE;
    if (!sor_missing(ns)) {
      code = COOP_NB_FOREIGN | ns;
      for (uint32_t j = 0; j < T; j++)
        if (sSlot[j] == ns) code = j;
    }
    sNb[q] = code;
  }
  __syncthreads();
  if (tid == 0u) {  // in (tile, face) order
    uint32_t k = 0;
    for (uint32_t q = 0; q < T * 6u; q++)
      if (sNb[q] != COOP_NB_NONE && (sNb[q] & COOP_NB_FOREIGN)) sFor[k++] = q;
    nFor = k;
  }
  for (uint32_t q = tid; q < T * 512u; q += COOP_WG) {
    uint32_t lt = q >> 9, l = q & 511u;
    sP[lt * 512u + coop_cidx(int32_t(l & 7u), int32_t((l >> 3) & 7u), int32_t(l >> 6))] = pres[sSlot[lt] * 512u + l];
  }
  for (uint32_t q = tid; q < 6u * 2u * 32u; q += COOP_WG) {
    const int3 a = coop_face_cell(q >> 6, q & 31u, (q >> 5) & 1u);
    CL.sFC[q] = uint16_t(coop_cidx(a.x, a.y, a.z));
  }
  __syncthreads();  // sFor and nFor
  // the exchange's addresses, fixed for the solve
  for (uint32_t k = tid; k < nFor; k += COOP_WG) {
    const uint32_t hf = sFor[k], lt = hf / 6u, f = hf - lt * 6u;
    CL.sPutL[k] = lt * 512u | (f << 16);
    CL.sPutG[k] = uint32_t(coop_halo_idx(0u, scap, sSlot[lt], f, 0u));
    CL.sGetG[k] = uint32_t(coop_halo_idx(0u, scap, sNb[hf] & ~COOP_NB_FOREIGN, f ^ 1u, 0u));
  }
  // Row items of each colour, those with a cell on a face toward another
  // workgroup's tile ("early") first: they are relaxed and sent before the
  // rest.  Within each part the rows inside the tile in y and z come before
  // the rows on its y or z faces, so that most waves hold rows of one kind and
  // skip the reads across those faces.  The lists are built in the halo copy's
  // space, filled only after.  Thread t classifies items 4t .. 4t + 3; an
  // exclusive scan of the four packed 16-bit counts places them.
  uint16_t* sList = reinterpret_cast<uint16_t*>(sH);  // [2][COOP_TMAX * 64]
  __shared__ uint64_t sWave[COOP_WG / 64];
  __shared__ uint32_t sEarly[2];
  const uint32_t nItems = T * 64u;
  for (uint32_t c = 0; c < 2u; c++) {
    uint32_t kinds = 0u;  // 2 bits per item: early ? 0 : 2, + 1 on a y or z face
    uint64_t cnt = 0u;
    for (uint32_t i = 0; i < 4u; i++) {
      const uint32_t q = tid * 4u + i;
      if (q >= nItems) continue;
      const uint32_t r = q & 63u, y = r & 7u, z = r >> 3;
      const uint32_t kind = (coop_row_early(sNb, q >> 6, r, c) ? 0u : 2u) + ((y == 0u || y == 7u || z == 0u || z == 7u) ? 1u : 0u);
      kinds |= kind << (2u * i);
      cnt += uint64_t(1) << (16u * kind);
    }
    uint64_t inc = cnt;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t lo = (uint32_t)__shfl_up((int)uint32_t(inc), off), hi = (uint32_t)__shfl_up((int)uint32_t(inc >> 32), off);
      if (int(lane_id()) >= off) inc += (uint64_t(hi) << 32) | lo;
    }
    if (lane_id() == 63u) sWave[tid >> 6] = inc;
    __syncthreads();
    uint64_t before = 0u, tot = 0u;
    for (uint32_t u = 0; u < COOP_WG / 64u; u++) {
      if (u < (tid >> 6)) before += sWave[u];
      tot += sWave[u];
    }
    const uint64_t ex = before + inc - cnt;
    uint32_t at[4], base = 0u;
    for (uint32_t kd = 0; kd < 4u; kd++) {
      at[kd] = base + uint32_t((ex >> (16u * kd)) & 0xffffu);
      base += uint32_t((tot >> (16u * kd)) & 0xffffu);
    }
    for (uint32_t i = 0; i < 4u; i++) {
      const uint32_t q = tid * 4u + i;
      if (q >= nItems) continue;
      sList[c * (COOP_TMAX * 64u) + at[(kinds >> (2u * i)) & 3u]++] = uint16_t(q);
    }
    if (tid == 0u) sEarly[c] = uint32_t(tot & 0xffffu) + uint32_t((tot >> 16) & 0xffffu);
    __syncthreads();
  }
  // this thread's items (positions tid + j COOP_WG of each list), with the rhs
  // and stencil bytes of their cells in registers for the whole solve
  uint32_t items[2][COOP_J];
  float4 rr[2][COOP_J];
  uint32_t rs[2][COOP_J];
#pragma unroll
  for (uint32_t c = 0; c < 2u; c++) {
#pragma unroll
    for (uint32_t j = 0; j < COOP_J; j++) {
      const uint32_t pos = tid + j * COOP_WG;
      items[c][j] = 0u;
      rr[c][j] = make_float4(0.f, 0.f, 0.f, 0.f);
      rs[c][j] = 0u;
      if (pos < nItems) {
        const uint32_t it = sList[c * (COOP_TMAX * 64u) + pos], r = it & 63u, y = r & 7u, z = r >> 3;
        const uint32_t base = sSlot[it >> 6] * 512u + 64u * z + 8u * y + ((y + z + c) & 1u);
        items[c][j] = it;
        rr[c][j] = make_float4(G.rhs[base], G.rhs[base + 2u], G.rhs[base + 4u], G.rhs[base + 6u]);
        rs[c][j] = uint32_t(G.sten[base]) | (uint32_t(G.sten[base + 2u]) << 8) | (uint32_t(G.sten[base + 4u]) << 16) |
                   (uint32_t(G.sten[base + 6u]) << 24);
      }
    }
  }
  const uint32_t nEarly[2] = {sEarly[0], sEarly[1]};
  __syncthreads();
  coop_zero_none(sH, sNb, T);
  coop_fill_start(pres, sH, sNb, sFor, nFor, 1u);
  __syncthreads();
  uint32_t epoch = 0;
  for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
    for (uint32_t c = 0; c < 2u; c++) {
      const bool last = it + 1u == iters && c == 1u;  // nothing more to exchange after it
      // what derives from the thread index, the items and the stencil bytes
      // (addresses, masks, divisors) is recomputed every half-sweep rather than
      // kept in registers for the whole solve: opaque copies
      uint32_t tv = tid;
      asm volatile("" : "+v"(tv));
#pragma unroll
      for (uint32_t j = 0; j < COOP_J; j++) asm volatile("" : "+v"(items[c][j]), "+v"(rs[c][j]));
      // early rows first
#pragma unroll
      for (uint32_t j = 0; j < COOP_J; j++)
        if (tv + j * COOP_WG < nEarly[c]) coop_relax_row(sP, sH, sNb, items[c][j], c, rr[c][j], rs[c][j], omega, ytab);
      __syncthreads();
      if (!last) {
        // colour-c face cells other workgroups read, tagged with this half-sweep
        ++epoch;
        for (uint32_t q = tid; q < nFor * 32u; q += COOP_WG) {
          const uint32_t k = q >> 5, e = q & 31u, pl = CL.sPutL[k];
          const float v = sP[(pl & 0xffffu) + CL.sFC[(((pl >> 16) << 1) + c) * 32u + e]];
          coop_halo_put(G, size_t(CL.sPutG[k]) + c * cStride + e, v, tag0 + epoch);
        }
      }
      // the other rows while the faces travel
#pragma unroll
      for (uint32_t j = 0; j < COOP_J; j++) {
        const uint32_t pos = tv + j * COOP_WG;
        if (pos >= nEarly[c] && pos < nItems) coop_relax_row(sP, sH, sNb, items[c][j], c, rr[c][j], rs[c][j], omega, ytab);
      }
      if (last) break;
      coop_fill_foreign(G, sH, sFor, CL.sGetG, nFor, c * cStride, tag0 + epoch, &sFail);
      __syncthreads();
      if (sFail) {  // a neighbour's faces never came: give up the solve (counted)
        if (tid == 0u) atomicAdd(&G.sorBar[COOP_MAX_WG + 1], 1u);
        return;
      }
    }
  }
  __syncthreads();
  for (uint32_t q = tid; q < T * 512u; q += COOP_WG) {
    uint32_t lt = q >> 9, l = q & 511u;
    pres[sSlot[lt] * 512u + l] = sP[lt * 512u + coop_cidx(int32_t(l & 7u), int32_t((l >> 3) & 7u), int32_t(l >> 6))];
  }
  if (w == 0u && tid == 0u) {
    G.sorBar[COOP_MAX_WG] = seq;
    G.sorBar[COOP_MAX_WG + 3] = 1u;
  }
}
__global__ void __launch_bounds__(COOP_WG) k_sor_resident(const DParams* __restrict__ PP, DGrid G, uint32_t iters,
                                                          uint32_t seq, uint32_t stream) {
  sor_resident_body(PP, G, iters, seq, stream, blockIdx.x, gridDim.x);
}

// Morton-ordered list of the slots in use (coopList) for the resident SOR
__global__ void k_coop_flags(const DParams* __restrict__ PP, DGrid G) {
  uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < PP->ntiles) G.coopFlags[m] = G.tmap[G.mortonTiles[m]] != ST_NONE ? 1u : 0u;
}
__global__ void k_coop_compact(const DParams* __restrict__ PP, DGrid G) {
  uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < PP->ntiles && G.coopFlags[m]) G.coopList[G.coopOffs[m]] = G.tmap[G.mortonTiles[m]];
}

// after an odd number of iterations the result sits in the partner buffer
__device__ __forceinline__ void copy_tiles_body(const DGrid& G, const float* __restrict__ src, float* __restrict__ dst,
                                                uint32_t seq, uint32_t bid, uint32_t nbid) {
  if (G.sorBar[COOP_MAX_WG] == seq) return;
  uint32_t tot = G.ctl[4] * 512u;
  for (uint32_t t = bid * blockDim.x + threadIdx.x; t < tot; t += nbid * blockDim.x) {
    uint32_t c = G.slotList[t >> 9] * 512u + (t & 511u);
    dst[c] = src[c];
  }
}
__global__ void k_copy_tiles(DGrid G, const float* __restrict__ src, float* __restrict__ dst, uint32_t seq) {
  copy_tiles_body(G, src, dst, seq, blockIdx.x, gridDim.x);
}

// projBody / project (wgsl-solver.js:165-180)
__device__ __forceinline__ void project_body(const DParams* __restrict__ PP, const DGrid& G, uint32_t bid, uint32_t nbid) {
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
__global__ void k_project(const DParams* __restrict__ PP, DGrid G) {
  project_body(PP, G, blockIdx.x, gridDim.x);
}

// Opt-in velocity extrapolation after the projection (no counterpart in the
// reference, which leaves faces without fluid as normFaces set them).  A mask
// byte per cell: bit a (axis a) marks a face whose value is known -- a face with
// fluid on one side, or one a layer has set -- and bit 3 + a a face that may be
// set, with air on both sides.  Each layer gives every settable face not yet
// known the mean of its known neighbours of the same axis (the faces of the six
// cells around, within the tile grid and in tiles with storage, summed in the
// order x-, x+, y-, y+, z-, z+), new and old values alike.  A layer reads the
// previous layer's mask and only faces known there, which it never writes, so
// the values are updated in place.
__device__ __forceinline__ void extrap_init_body(const DParams* __restrict__ PP, const DGrid& G, uint8_t* __restrict__ mask,
                                                 uint32_t bid, uint32_t nbid) {
  const DParams& P = *PP;
  FOR_ACTIVE(G) {
    const uint32_t s = G.slotList[t >> 9], local = t & 511u, c = s * 512u + local;
    const int3 at = d_slotCoord(P, G.slotTile, s, local);
    const uint32_t th = d_ctypeOf(P, G.tmap, G.ctype, at.x, at.y, at.z);
    uint32_t m = 0u;
    for (uint32_t axis = 0; axis < 3; axis++) {
      int3 lo = at;
      if (axis == 0) lo.x -= 1; else if (axis == 1) lo.y -= 1; else lo.z -= 1;
      const uint32_t tl = d_ctypeOf(P, G.tmap, G.ctype, lo.x, lo.y, lo.z);
      if (tl == 1u || th == 1u) m |= 1u << axis;
      else if (tl == 0u && th == 0u) m |= 8u << axis;
    }
    mask[c] = uint8_t(m);
  }
}
__global__ void k_extrap_init(const DParams* __restrict__ PP, DGrid G, uint8_t* __restrict__ mask) {
  extrap_init_body(PP, G, mask, blockIdx.x, gridDim.x);
}
__device__ __forceinline__ void extrap_layer_body(const DParams* __restrict__ PP, const DGrid& G, const uint8_t* __restrict__ min,
                                                  uint8_t* __restrict__ mout, uint32_t bid, uint32_t nbid) {
  const DParams& P = *PP;
  const int32_t gx = int32_t(P.tx) * 8, gy = int32_t(P.ty) * 8, gz = int32_t(P.tz) * 8;
  FOR_ACTIVE(G) {
    const uint32_t s = G.slotList[t >> 9], local = t & 511u, c = s * 512u + local;
    const int3 at = d_slotCoord(P, G.slotTile, s, local);
    const uint32_t m = min[c];
    uint32_t mo = m;
    for (uint32_t axis = 0; axis < 3; axis++) {
      if (((m >> axis) & 1u) || !((m >> (3u + axis)) & 1u)) continue;
      float sn = 0.0f, so = 0.0f;
      uint32_t cnt = 0u;
      for (uint32_t q = 0; q < 6u; q++) {
        int3 r = at;
        const int32_t d = (q & 1u) ? 1 : -1;
        if (q < 2u) r.x += d; else if (q < 4u) r.y += d; else r.z += d;
        if (r.x < 0 || r.y < 0 || r.z < 0 || r.x >= gx || r.y >= gy || r.z >= gz) continue;
        const uint32_t sl = G.tmap[d_tileOf(P, r.x, r.y, r.z)];
        if (sl == ST_NONE) continue;
        const uint32_t cn = sl * 512u + uint32_t(r.x & 7) + 8u * (uint32_t(r.y & 7) + 8u * uint32_t(r.z & 7));
        if (!((min[cn] >> axis) & 1u)) continue;
        sn = sn + gvR(G, cn * 3u + axis);
        so = so + gv0R(G, cn * 3u + axis);
        cnt++;
      }
      if (cnt) {
        const uint32_t f = c * 3u + axis;
        G.accV[f] = (int32_t)st_f2u(sn / float(cnt));
        G.accW[f] = (int32_t)st_f2u(so / float(cnt));
        mo |= 1u << axis;
      }
    }
    mout[c] = uint8_t(mo);
  }
}
__global__ void k_extrap_layer(const DParams* __restrict__ PP, DGrid G, const uint8_t* __restrict__ min, uint8_t* __restrict__ mout) {
  extrap_layer_body(PP, G, min, mout, blockIdx.x, gridDim.x);
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

// Grid access for g2p: from the global arrays (tmap lookups), or from a tile
// window staged in LDS.  Both return the same words.
struct GlobalGrid;
struct WindowGrid;
__device__ __forceinline__ void d_sampleBothWin(const DParams& P, const WindowGrid& g, float3 p, float3& vn, float3& vo);
template <class GridT>
__device__ __forceinline__ void d_sampleBoth(const DParams& P, const GridT& g, float3 p, float3& vn, float3& vo);
struct GlobalGrid {
  const DParams& P;
  const DGrid& G;
  __device__ void sample(float3 p, float3& vn, float3& vo) const { d_sampleBoth(P, *this, p, vn, vo); }
  __device__ float fnew(uint32_t axis, int32_t i, int32_t j, int32_t k) const { return gvR(G, d_faceIndex(P, G.tmap, axis, i, j, k)); }
  __device__ float fold(uint32_t axis, int32_t i, int32_t j, int32_t k) const { return gv0R(G, d_faceIndex(P, G.tmap, axis, i, j, k)); }
  __device__ int32_t cnt(int32_t i, int32_t j, int32_t k) const { return G.cnt[d_cellIdx(P, G.tmap, i, j, k)]; }
  // the particle counts of the 2x2x2 cells from bc, each clamped to the box, summed
  __device__ float dsum(int3 bc) const {
    float csum = 0.0f;
    for (int32_t dz = 0; dz <= 1; dz++)
      for (int32_t dy = 0; dy <= 1; dy++)
        for (int32_t dx = 0; dx <= 1; dx++) {
          int32_t cx = d_iclamp(bc.x + dx, 0, int32_t(P.nx) - 1), cy = d_iclamp(bc.y + dy, 0, int32_t(P.ny) - 1),
                  cz = d_iclamp(bc.z + dz, 0, int32_t(P.nz) - 1);
          csum = csum + float(cnt(cx, cy, cz));
        }
    return csum;
  }
};
// The 2x2x2 count sums of a tile window, for every base cell a particle of the
// tile can have: base b = org + w, w in [0, 8] per axis, cells clamped to the box
// as GlobalGrid::dsum clamps them.  Counts are integers far below 2^24, so their
// f32 sum in any order is exact and equals the integer sum.
#define DSUM_N 9
#define DSUM3 (DSUM_N * DSUM_N * DSUM_N)
struct WindowGrid {
  const DParams& P;
  const DGrid& G;
  const float2 (*sNO)[WIN3];  // per axis and window cell: (new, old) face velocity
  const int32_t* sC;
  int3 org;
  __device__ void sample(float3 p, float3& vn, float3& vo) const { d_sampleBothWin(P, *this, p, vn, vo); }
  // cells of the density sum come clamped to the box, so they lie in the window
  __device__ int32_t cnt(int32_t i, int32_t j, int32_t k) const {
    return sC[win_idx(wclamp(i - org.x), wclamp(j - org.y), wclamp(k - org.z))];
  }
  // GlobalGrid::dsum from the table sS.  A base below -1 (above n - 1) clamps its
  // two cells to the ones of base -1 (n - 1), so the base is clamped to
  // [-1, n - 1] first; for a particle of the tile that puts it in the window's
  // [0, 8] (the outer clamp only keeps LDS reads in bounds)
  const int32_t* sS;
  __device__ float dsum(int3 bc) const {
    const int32_t wx = min(max(d_iclamp(bc.x, -1, int32_t(P.nx) - 1) - org.x, 0), DSUM_N - 1);
    const int32_t wy = min(max(d_iclamp(bc.y, -1, int32_t(P.ny) - 1) - org.y, 0), DSUM_N - 1);
    const int32_t wz = min(max(d_iclamp(bc.z, -1, int32_t(P.nz) - 1) - org.z, 0), DSUM_N - 1);
    return float(sS[uint32_t(wx) + __umul24(uint32_t(wy), uint32_t(DSUM_N)) + __umul24(uint32_t(wz), uint32_t(DSUM_N * DSUM_N))]);
  }
};

// sampleBoth (wgsl-core.js:138-154) from a staged window, with the six
// per-dimension terms of win_axes; each stencil's arithmetic is the one of
// d_sampleBoth, term for term.
__device__ __forceinline__ void d_sampleBothWin(const DParams& P, const WindowGrid& g, float3 p, float3& vn, float3& vo) {
  const WinAxes wa = win_axes(P, p, g.org);
  float rn[3], ro[3];
#pragma unroll
  for (uint32_t axis = 0; axis < 3; axis++) {
    const WinAxis& ax = wa.x(axis);
    const WinAxis& ay = wa.y(axis);
    const WinAxis& az = wa.z(axis);
    const float2* __restrict__ s = g.sNO[axis];
    float sn = 0.0f, so = 0.0f;
#pragma unroll
    for (uint32_t c = 0; c < 8; c++) {
      const uint32_t dx = c & 1u, dy = (c >> 1) & 1u, dz = (c >> 2) & 1u;
      const float w = (dx ? ax.f : ax.g) * (dy ? ay.f : ay.g) * (dz ? az.f : az.g);
      const float2 e = s[win_idx(ax.w[dx], ay.w[dy], az.w[dz])];
      sn = sn + w * e.x;
      so = so + w * e.y;
    }
    rn[axis] = sn;
    ro[axis] = so;
  }
  vn = make_float3(rn[0], rn[1], rn[2]);
  vo = make_float3(ro[0], ro[1], ro[2]);
}

// sampleBoth (wgsl-core.js:138-154)
template <class GridT>
__device__ __forceinline__ void d_sampleBoth(const DParams& P, const GridT& g, float3 p, float3& vn, float3& vo) {
  float rn[3], ro[3];
  for (uint32_t axis = 0; axis < 3; axis++) {
    float3 q = make_float3((p.x - P.x0) * P.ih, (p.y - P.y0) * P.ih, (p.z - P.z0) * P.ih);
    if (axis == 0) { q.y = q.y - 0.5f; q.z = q.z - 0.5f; }
    else if (axis == 1) { q.x = q.x - 0.5f; q.z = q.z - 0.5f; }
    else { q.x = q.x - 0.5f; q.y = q.y - 0.5f; }
    float3 f0 = make_float3(floorf(q.x), floorf(q.y), floorf(q.z));
    float3 fr = f3sub(q, f0);
    int3 i0 = make_int3(st_f2i_sat(f0.x), st_f2i_sat(f0.y), st_f2i_sat(f0.z));
    float sn = 0.0f, so = 0.0f;
    for (uint32_t c = 0; c < 8; c++) {
      int32_t dx = int32_t(c & 1u), dy = int32_t((c >> 1) & 1u), dz = int32_t((c >> 2) & 1u);
      float w = (dx == 1 ? fr.x : 1.0f - fr.x) * (dy == 1 ? fr.y : 1.0f - fr.y) * (dz == 1 ? fr.z : 1.0f - fr.z);
      sn = sn + w * g.fnew(axis, i0.x + dx, i0.y + dy, i0.z + dz);
      so = so + w * g.fold(axis, i0.x + dx, i0.y + dy, i0.z + dz);
    }
    rn[axis] = sn;
    ro[axis] = so;
  }
  vn = make_float3(rn[0], rn[1], rn[2]);
  vo = make_float3(ro[0], ro[1], ro[2]);
}

// one particle of g2p: FLIP/PIC blend, foam, air, speed cap, Euler advection,
// collisions, open faces (wgsl-solver.js:208-275)
// The part of g2p that depends only on the density sum csum (the count of the
// 8 cells around the particle, or 8 x its own cell's count): air, shield, the
// foam decay exp(-dt / tau) and the drag factor 1 - exp(-kd dt)
// (wgsl-solver.js:221-245).  Per step, k_foam_table evaluates it for every
// integer csum below FOAM_TAB; g2p reads the table, which gives the same bits
// as evaluating it in place.
__device__ __forceinline__ float4 d_foamRow(const DParams& P, float csum) {
  float rho = csum / (8.0f * P.rest);
  float air = 1.0f - st_smoothstepf(0.008f, 0.1f, rho);
  float shield = st_smoothstepf(0.1f, 0.65f, rho);
  float exposure = 1.0f - 0.97f * shield;
  float tau = fmaxf(P.q5, 0.05f) * st_mixf(1.0f, 0.35f, st_smoothstepf(0.7f, 1.0f, rho)) * st_mixf(1.0f, 2.5f, air);
  float kd = P.q4 * (0.04f + 1.5f * exposure * exposure);
  return make_float4(air, shield, st_expf(-P.dt / tau), 1.0f - st_expf(-kd * P.dt));
}

__device__ __forceinline__ void foam_table_body(const DParams* __restrict__ PP, const DGrid& G, uint32_t bid, uint32_t nbid) {
  for (uint32_t t = bid * blockDim.x + threadIdx.x; t < FOAM_TAB; t += nbid * blockDim.x)
    G.foamTab[t] = d_foamRow(*PP, float(t));
}
__global__ void k_foam_table(const DParams* __restrict__ PP, DGrid G) {
  foam_table_body(PP, G, blockIdx.x, gridDim.x);
}

template <class GridT>
__device__ __forceinline__ void g2p_particle(const DParams& P, const DCols* __restrict__ C, const DGrid& G, const GridT& g,
                                             bool storage, float4 p4, float4 v04, float4& po, float4& vo4) {
  const float* __restrict__ sdf = G.sdf;
  float3 p = make_float3(p4.x, p4.y, p4.z);
  float3 v0 = make_float3(v04.x, v04.y, v04.z);
  int3 fcell = d_particleCell(P, p);
  float3 vn = make_float3(v0.x - 0.0f, v0.y - P.gravity * P.dt, v0.z - 0.0f);
  float3 vo = v0;
  if (storage) g.sample(p, vn, vo);
  float3 fl = f3add(v0, f3sub(vn, vo));
  float3 v = make_float3(st_mixf(vn.x, fl.x, P.flip), st_mixf(vn.y, fl.y, P.flip), st_mixf(vn.z, fl.z, P.flip));
  float lw = 0.0f;
  if (P.ocean != 0u) {  // near the ocean's boundary, ease into the waves (wgsl-solver.js:218)
    lw = d_layerW(P, p);
    if (lw > 0.0f) {
      const float s = st_clampf(lw * 2.0f, 0.0f, 1.0f);
      const float3 ov = d_oceanVel(P, p);
      const float t = s * s * P.oease;
      v = make_float3(st_mixf(v.x, ov.x, t), st_mixf(v.y, ov.y, t), st_mixf(v.z, ov.z, t));
    }
  }
  int3 bc = make_int3(st_f2i_sat(floorf((p.x - P.x0) * P.ih - 0.5f)), st_f2i_sat(floorf((p.y - P.y0) * P.ih - 0.5f)), st_f2i_sat(floorf((p.z - P.z0) * P.ih - 0.5f)));
  float csum = 8.0f * P.rest;
  bool fast = P.q12 > 0.5f;
  if (P.q9 < 0.5f && fast) {
    csum = 8.0f * float(g.cnt(fcell.x, fcell.y, fcell.z));
  } else if (P.q9 < 0.5f) {
    csum = g.dsum(bc);
  }
  // air, shield, foam decay and drag factor depend on the density sum alone
  const float4 ft = (P.q9 < 0.5f && csum >= 0.0f && csum < float(FOAM_TAB)) ? G.foamTab[int32_t(csum)] : d_foamRow(P, csum);
  const float air = ft.x, shield = ft.y;
  float sens = fmaxf(P.q6, 0.05f);
  const float settle = P.settle;
  // impact and turb are +0 wherever the quotient they clamp is not above 15
  // (0.35); below the squared lengths impS2 (turbS2) it cannot be, so the
  // roots and divisions are skipped there with the same result
  float impact = 0.0f, turb = 0.0f;
  const float3 dvp = f3sub(vn, vo), dvt = f3sub(v0, vo);
  if (f3dot(dvp, dvp) > P.impS2) {
    float pacc = f3len(dvp) / P.dt;
    impact = st_clampf((pacc / sens - 15.0f) / 80.0f, 0.0f, 1.0f) * settle;
  }
  if (f3dot(dvt, dvt) > P.turbS2) turb = st_clampf((f3len(dvt) / sens - 0.35f) / 1.5f, 0.0f, 1.0f) * settle;
  // +0 where air is +0 (the factor it scales is finite), without the root
  float spray = 0.0f;
  if (air != 0.0f) spray = air * (0.25f + 0.75f * st_clampf(f3len(v) / 4.0f, 0.0f, 1.0f));
  float foam = v04.w * ft.z;
  foam = fmaxf(foam, fmaxf(spray * 0.9f, fmaxf(impact, turb * 0.8f)));
  if (P.ocean != 0u) foam = foam * (1.0f - st_clampf(lw * 2.0f, 0.0f, 1.0f));  // easing is not churning (wgsl-solver.js:248)
  float3 gust = make_float3(0.0f, 0.0f, 0.0f);
  if (P.q7 > 0.0f && P.q4 > 0.0f && (!fast || shield < 0.9f)) {
    float3 gp = f3mul(p, 0.55f);
    float tt = P.q8 * 0.6f;
    gust = f3sub(f3mul(make_float3(d_vnz(f3add(gp, make_float3(tt, 0.0f, 0.0f))), d_vnz(f3add(gp, make_float3(17.3f, tt, 4.1f))),
                                   d_vnz(f3add(gp, make_float3(5.7f, 9.2f, tt)))), 2.0f),
                 make_float3(1.0f, 1.0f, 1.0f));
  }
  float3 airv = f3add(make_float3(P.q1, P.q2, P.q3), f3mul(gust, P.q7));
  v = f3add(v, f3mul(f3sub(airv, v), ft.w));
  float sp = f3len(v);
  bool spIsLen = true;  // sp still equals |v| (v untouched since)
  if (sp > P.vmax) {
    v = f3mul(v, P.vmax / sp);
    spIsLen = false;
  }
  float3 pt = f3add(p, f3mul(v, P.dt));
  if (P.rk2 != 0u && storage) {
    // opt-in midpoint rule through the grid's new velocity u (vn is u at p); the
    // midpoint may lie outside the tile's window, so it is sampled from the grid
    const float hdt = 0.5f * P.dt;
    const float3 pm = make_float3(p.x + vn.x * hdt, p.y + vn.y * hdt, p.z + vn.z * hdt);
    float3 um, uo;
    GlobalGrid gm{P, G};
    d_sampleBoth(P, gm, pm, um, uo);
    pt = make_float3(p.x + um.x * P.dt, p.y + um.y * P.dt, p.z + um.z * P.dt);
  }
  float3 pc = d_collideP(P, C, sdf, pt);
  float3 dsp = f3sub(pc, pt);
  if (f3dot(dsp, dsp) > 1.0e-12f) {
    float3 nrm = f3div(dsp, f3len(dsp));
    float vd = f3dot(v, nrm);
    if (vd < 0.0f) {
      v = f3sub(v, f3mul(nrm, vd));
      spIsLen = false;
    }
  }
  uint32_t o = P.open;
  if (o != 0u && (((o & 1u) && pc.x < P.x0) || ((o & 2u) && pc.x > P.x1) || ((o & 4u) && pc.y < P.y0) || ((o & 8u) && pc.y > P.y1) ||
                  ((o & 16u) && pc.z < P.z0) || ((o & 32u) && pc.z > P.z1))) {
    po = make_float4(0.0f, -1.0e6f, 0.0f, 0.0f);
    vo4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  float vlen = sp;
  if (!spIsLen) vlen = f3len(v);
  po = make_float4(pc.x, pc.y, pc.z, fminf(vlen, 63.0f) + 64.0f * floorf(foam * 15.0f + 0.5f));
  vo4 = make_float4(v.x, v.y, v.z, foam);
}

// g2p over the bins: one workgroup per tile (bin ntiles holds the dead
// particles, which are copied).  Reads A through the bin list and writes the new
// state into B in bin order; the host then swaps A and B.
// With `keys` set it also bins the new positions for the next substep
// (bin_mark); the particle loops then run whole waves, so every lane calls it.
__device__ __forceinline__ void g2p_emit(const DParams& P, const DGrid& G, uint32_t keys, uint32_t k, uint32_t i, bool valid,
                                         float4 po, float4 vo) {
  if (valid) {
    G.posB[k] = po;
    G.velB[k] = vo;
    G.pidB[k] = G.pid[i];
  }
  if (keys) bin_mark(P, G, k, valid, po);
}

__device__ __forceinline__ void g2p_tile_body(const DParams* __restrict__ PP, const DCols* __restrict__ C, const DGrid& G,
                                              uint32_t keys, uint32_t bid, uint32_t nbid) {
  const DParams& P = *PP;
  __shared__ float2 sNO[3][WIN3];
  __shared__ int32_t sC[WIN3];
  __shared__ int32_t sS[DSUM3];
  __shared__ uint32_t wslots[27];
  for (uint32_t t = bid; t <= P.ntiles; t += nbid) {
    uint32_t beg = G.binStart[t], end = G.binStart[t + 1];
    if (beg == end) continue;
    if (t == P.ntiles) {
      for (uint32_t k0 = beg; k0 < end; k0 += WG) {
        uint32_t k = k0 + threadIdx.x;
        bool valid = k < end;
        uint32_t i = valid ? G.binIdx[k] : 0u;
        float4 po = make_float4(0.f, -1.0e6f, 0.f, 0.f), vo = po;
        if (valid) {
          po = G.pos[i];
          vo = G.vel[i];
        }
        g2p_emit(P, G, keys, k, i, valid, po, vo);
      }
      continue;
    }
    bool storage = G.tmap[t] != ST_NONE;
    if (!storage) {  // its block had no room this step: falls freely (wgsl-solver.js:215)
      GlobalGrid gg{P, G};
      for (uint32_t k0 = beg; k0 < end; k0 += WG) {
        uint32_t k = k0 + threadIdx.x;
        bool valid = k < end;
        uint32_t i = valid ? G.binIdx[k] : 0u;
        float4 po = make_float4(0.f, -1.0e6f, 0.f, 0.f), vo = po;
        if (valid) g2p_particle(P, C, G, gg, false, G.pos[i], G.vel[i], po, vo);
        g2p_emit(P, G, keys, k, i, valid, po, vo);
      }
      continue;
    }
    int3 org;
    load_wslots(P, G, G.tmap[t], t, wslots, org);
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < WIN3; q += WG) {
      int32_t wx = int32_t(q % WIN), wy = int32_t((q / WIN) % WIN), wz = int32_t(q / (WIN * WIN));
      uint32_t c = wcell(wslots, wx, wy, wz, org);
      if (c == ST_NONE) {
        sNO[0][q] = make_float2(0.0f, 0.0f); sNO[1][q] = make_float2(0.0f, 0.0f); sNO[2][q] = make_float2(0.0f, 0.0f);
        sC[q] = 0;
        continue;
      }
      sNO[0][q] = make_float2(gvR(G, c * 3), gv0R(G, c * 3));
      sNO[1][q] = make_float2(gvR(G, c * 3 + 1), gv0R(G, c * 3 + 1));
      sNO[2][q] = make_float2(gvR(G, c * 3 + 2), gv0R(G, c * 3 + 2));
      sC[q] = G.cnt[c];
    }
    __syncthreads();
    if (P.q9 < 0.5f && !(P.q12 > 0.5f)) {  // the density sums g2p reads
      for (uint32_t q = threadIdx.x; q < DSUM3; q += WG) {
        const int32_t bx = org.x + int32_t(q % DSUM_N), by = org.y + int32_t((q / DSUM_N) % DSUM_N), bz = org.z + int32_t(q / (DSUM_N * DSUM_N));
        int32_t n = 0;
        for (int32_t dz = 0; dz <= 1; dz++)
          for (int32_t dy = 0; dy <= 1; dy++)
            for (int32_t dx = 0; dx <= 1; dx++) {
              const int32_t cx = d_iclamp(bx + dx, 0, int32_t(P.nx) - 1), cy = d_iclamp(by + dy, 0, int32_t(P.ny) - 1),
                            cz = d_iclamp(bz + dz, 0, int32_t(P.nz) - 1);
              n += sC[wclamp(cx - org.x) + WIN * (wclamp(cy - org.y) + WIN * wclamp(cz - org.z))];
            }
        sS[q] = n;
      }
      __syncthreads();
    }
    WindowGrid wg{P, G, sNO, sC, org, sS};
    for (uint32_t k0 = beg; k0 < end; k0 += WG) {
      uint32_t k = k0 + threadIdx.x;
      bool valid = k < end;
      uint32_t i = valid ? G.binIdx[k] : 0u;
      float4 po = make_float4(0.f, -1.0e6f, 0.f, 0.f), vo = po;
      if (valid) g2p_particle(P, C, G, wg, true, G.pos[i], G.vel[i], po, vo);
      g2p_emit(P, G, keys, k, i, valid, po, vo);
    }
    __syncthreads();
  }
}
__global__ void __launch_bounds__(WG) k_g2p_tile(const DParams* __restrict__ PP, const DCols* __restrict__ C, DGrid G,
                                                 uint32_t keys) {
  g2p_tile_body(PP, C, G, keys, blockIdx.x, gridDim.x);
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
// narrow band (wgsl-narrowband.js, wgsl-core.js:156-238), with the ocean's
// branches (the ocean turns the narrow band on)

// signed distance in cells to the starting water: boxes, spheres, water volumes and
// the ocean (wgsl-core.js:228-238)
__device__ __forceinline__ float d_fillPhi(const DParams& P, const DCols* __restrict__ C, const float* __restrict__ sdf, float3 x) {
  float dd = 1.0e9f;
  for (uint32_t k = 0; k < P.nfill; k++) {
    const float* A = P.fills[k * 2];
    const float* B = P.fills[k * 2 + 1];
    if (A[0] < 1.5f) {
      float3 q = make_float3(fmaxf(A[1] - x.x, x.x - B[0]), fmaxf(A[2] - x.y, x.y - B[1]), fmaxf(A[3] - x.z, x.z - B[2]));
      float3 qp = make_float3(fmaxf(q.x, 0.0f), fmaxf(q.y, 0.0f), fmaxf(q.z, 0.0f));
      dd = fminf(dd, f3len(qp) + fminf(fmaxf(q.x, fmaxf(q.y, q.z)), 0.0f));
    } else {
      dd = fminf(dd, f3len(f3sub(x, make_float3(A[1], A[2], A[3]))) - B[0]);
    }
  }
  for (uint32_t k = P.ncols; k < P.ncols + P.nfv; k++) {
    float phi;
    float3 wn;
    d_sdfEval(C, sdf, k, x, phi, wn);
    dd = fminf(dd, phi);
  }
  if (P.ocean != 0u) dd = fminf(dd, x.y - d_oceanEta(P, x.x, x.z));
  return dd * P.ih;
}
// depth at a point from the cell centres around it, collider cells left out
// (wgsl-core.js:169-183)
__device__ __forceinline__ float d_phiSample(const DParams& P, const DGrid& G, float3 x) {
  float3 g = make_float3((x.x - P.x0) * P.ih - 0.5f, (x.y - P.y0) * P.ih - 0.5f, (x.z - P.z0) * P.ih - 0.5f);
  float3 gc = make_float3(st_clampf(g.x, 0.0f, float(P.nx) - 1.001f), st_clampf(g.y, 0.0f, float(P.ny) - 1.001f),
                          st_clampf(g.z, 0.0f, float(P.nz) - 1.001f));
  float3 f0 = make_float3(floorf(gc.x), floorf(gc.y), floorf(gc.z));
  float3 fr = f3sub(gc, f0);
  int3 i0 = make_int3(st_f2i_sat(f0.x), st_f2i_sat(f0.y), st_f2i_sat(f0.z));
  float s = 0.0f, ws = 0.0f;
  for (uint32_t c = 0; c < 8; c++) {
    int32_t dx = int32_t(c & 1u), dy = int32_t((c >> 1) & 1u), dz = int32_t((c >> 2) & 1u);
    float w = (dx == 1 ? fr.x : 1.0f - fr.x) * (dy == 1 ? fr.y : 1.0f - fr.y) * (dz == 1 ? fr.z : 1.0f - fr.z);
    uint32_t ct = G.ctype[d_cellIdx(P, G.tmap, i0.x + dx, i0.y + dy, i0.z + dz)];
    if (((ct >> 8) & 0xffu) != 0u || w <= 0.0f) continue;
    s = s + w * d_phiOfBits(ct);
    ws = ws + w;
  }
  if (ws < 1.0e-4f) return 0.5f;
  return s / ws;
}

// tileMarkPhi (wgsl-narrowband.js:6-10): last step's tiles holding water by depth
__device__ __forceinline__ void nb_mark_phi_body(const DParams* __restrict__ PP, const DGrid& G, uint32_t bid, uint32_t nbid) {
  const DParams& P = *PP;
  FOR_ACTIVE(G) {
    uint32_t s = G.slotList[t >> 9], local = t & 511u;
    int3 cc = d_slotCoord(P, G.slotTile, s, local);
    if (!d_inDom(P, cc) || d_phiOfBits(G.ctype[s * 512u + local]) >= 0.5f) continue;
    int32_t lx = max(cc.x - 2, 0) >> 3, ly = max(cc.y - 2, 0) >> 3, lz = max(cc.z - 2, 0) >> 3;
    int32_t hx = min(cc.x + 2, int32_t(P.tx) * 8 - 1) >> 3, hy = min(cc.y + 2, int32_t(P.ty) * 8 - 1) >> 3,
            hz = min(cc.z + 2, int32_t(P.tz) * 8 - 1) >> 3;
    for (int32_t tk = lz; tk <= hz; tk++)
      for (int32_t tj = ly; tj <= hy; tj++)
        for (int32_t ti = lx; ti <= hx; ti++) G.need[uint32_t(ti) + P.tx * (uint32_t(tj) + P.ty * uint32_t(tk))] = 1u;
  }
}
__global__ void k_nb_mark_phi(const DParams* __restrict__ PP, DGrid G) {
  nb_mark_phi_body(PP, G, blockIdx.x, gridDim.x);
}
// tileMarkFill (wgsl-narrowband.js:13-18): on a start, every tile the starting water reaches
__global__ void k_nb_mark_fill(const DParams* __restrict__ PP, const DCols* __restrict__ C, DGrid G) {
  const DParams& P = *PP;
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P.ntiles) return;
  uint32_t tt = t / P.tx;
  float3 tc = make_float3(float(t % P.tx), float(tt % P.ty), float(tt / P.ty));
  tc = f3add(f3mul(tc, 8.0f), make_float3(4.0f, 4.0f, 4.0f));
  float3 q = f3add(make_float3(P.x0, P.y0, P.z0), f3mul(tc, P.h));
  if (d_fillPhi(P, C, G.sdf, q) < 9.0f) G.need[t] = 1u;
}
// phiInitAll (wgsl-narrowband.js:21-35)
__device__ __forceinline__ void nb_phi_init_body(const DParams* __restrict__ PP, const DCols* __restrict__ C,
                                                 const DGrid& G, uint32_t bid, uint32_t nbid) {
  const DParams& P = *PP;
  FOR_ACTIVE(G) {
    uint32_t s = G.slotList[t >> 9], local = t & 511u, c = s * 512u + local;
    int3 p = d_slotCoord(P, G.slotTile, s, local);
    float ph = st_clampf(d_fillPhi(P, C, G.sdf, d_cellCenter(P, p.x, p.y, p.z)), -7.9f, 7.9f);
    if (ph < 0.0f) ph = fminf(ph, -1.5f);
    if (!d_inDom(P, p)) ph = 8.0f;
    uint32_t ct = G.ctype[c];
    if (P.nbInit == 2u) ph = fminf(ph, d_phiOfBits(ct));
    G.ctype[c] = d_phiPutBits(ct, ph);
    if (P.nbInit == 1u) {  // at rest, or moving with the waves under the ocean's surface
      for (uint32_t ax = 0; ax < 3; ax++) {
        float v = 0.0f;
        if (P.ocean != 0u) {
          float3 fp = d_facePos(P, p, ax);
          if (fp.y < d_oceanEta(P, fp.x, fp.z)) v = f3get(d_oceanVel(P, fp), ax);
        }
        G.velG[c * 3 + ax] = v;
      }
    }
  }
}
__global__ void k_nb_phi_init(const DParams* __restrict__ PP, const DCols* __restrict__ C, DGrid G) {
  nb_phi_init_body(PP, C, G, blockIdx.x, gridDim.x);
}
// nbPhi (wgsl-narrowband.js:39-49): depth carried by the kept grid velocity
__device__ __forceinline__ void nb_phi_body(const DParams* __restrict__ PP, const DGrid& G, uint32_t bid, uint32_t nbid) {
  const DParams& P = *PP;
  FOR_ACTIVE(G) {
    uint32_t s = G.slotList[t >> 9], local = t & 511u, c = s * 512u + local;
    int3 p = d_slotCoord(P, G.slotTile, s, local);
    if (!d_inDom(P, p)) { G.rhs[c] = 8.0f; continue; }
    if (((G.ctype[c] >> 8) & 0xffu) != 0u) { G.rhs[c] = 0.5f; continue; }
    float3 x = d_cellCenter(P, p.x, p.y, p.z);
    float phA = d_phiSample(P, G, f3sub(x, f3mul(d_velGVec(P, G, x), P.dt)));
    bool inside = G.cnt[c] > 0 || phA < -1.0f;
    if (P.ocean != 0u && d_layerW(P, x) >= 0.5f) inside = x.y < d_oceanEta(P, x.x, x.z);  // the ocean surface decides
    G.rhs[c] = inside ? st_clampf(phA, -7.9f, -0.5f) : st_clampf(phA, 0.5f, 7.9f);
  }
}
__global__ void k_nb_phi(const DParams* __restrict__ PP, DGrid G) {
  nb_phi_body(PP, G, blockIdx.x, gridDim.x);
}
// jacNb, nbJacA (fromRhs: phi <- from rhs) and nbJacB (rhs <- from phi) (wgsl-narrowband.js:51-68)
__device__ __forceinline__ void nb_jac_body(const DParams* __restrict__ PP, const DGrid& G, uint32_t fromRhs, uint32_t bid,
                                            uint32_t nbid) {
  const DParams& P = *PP;
  FOR_ACTIVE(G) {
    uint32_t s = G.slotList[t >> 9], local = t & 511u, c = s * 512u + local;
    int3 p = d_slotCoord(P, G.slotTile, s, local);
    float cur = fromRhs ? G.rhs[c] : d_phiOfBits(G.ctype[c]);
    float out;
    if (!d_inDom(P, p)) out = 8.0f;
    else if (cur >= 0.0f) out = cur;
    else {
      float m = -8.0f;
      for (uint32_t q = 0; q < 6; q++) {
        int3 n = p;
        if (q == 0) n.x -= 1; else if (q == 1) n.x += 1; else if (q == 2) n.y -= 1; else if (q == 3) n.y += 1; else if (q == 4) n.z -= 1; else n.z += 1;
        if (!d_inDom(P, n)) continue;
        uint32_t cq = d_cellIdx(P, G.tmap, n.x, n.y, n.z);
        uint32_t ct = G.ctype[cq];
        if (((ct >> 8) & 0xffu) != 0u) continue;
        m = fmaxf(m, fromRhs ? G.rhs[cq] : d_phiOfBits(ct));
      }
      out = st_clampf(m - 1.0f, -7.9f, -0.5f);
    }
    if (fromRhs) G.ctype[c] = d_phiPutBits(G.ctype[c], out);
    else G.rhs[c] = out;
  }
}
__global__ void k_nb_jac(const DParams* __restrict__ PP, DGrid G, uint32_t fromRhs) {
  nb_jac_body(PP, G, fromRhs, blockIdx.x, gridDim.x);
}
// nbVel (wgsl-narrowband.js:71-85): water faces no particle reached take the kept
// grid velocity, carried along
__device__ __forceinline__ void nb_vel_body(const DParams* __restrict__ PP, const DGrid& G, uint32_t bid, uint32_t nbid) {
  const DParams& P = *PP;
  FOR_ACTIVE(G) {
    uint32_t s = G.slotList[t >> 9], local = t & 511u, c = s * 512u + local;
    int3 p = d_slotCoord(P, G.slotTile, s, local);
    if (!d_inDom(P, p) || (G.ctype[c] & 0xffu) != 1u) continue;
    for (uint32_t ax = 0; ax < 3; ax++) {
      int3 lo = p;
      if (ax == 0) lo.x -= 1; else if (ax == 1) lo.y -= 1; else lo.z -= 1;
      if (!d_inDom(P, lo)) continue;
      if ((G.ctype[d_cellIdx(P, G.tmap, lo.x, lo.y, lo.z)] & 0xffu) != 1u) continue;
      uint32_t f = c * 3u + ax;
      if (G.accW[f] > 0) continue;
      float3 fp = d_facePos(P, p, ax);
      float vc = d_velGAxis(P, G, f3sub(fp, f3mul(d_velGVec(P, G, fp), P.dt)), ax);
      G.accV[f] = st_f2i_sat(rintf(vc * P.scale));
      G.accW[f] = st_f2i_sat(P.scale);
    }
  }
}
__global__ void k_nb_vel(const DParams* __restrict__ PP, DGrid G) {
  nb_vel_body(PP, G, blockIdx.x, gridDim.x);
}
// saveVel (wgsl-narrowband.js:88-91)
__device__ __forceinline__ void save_vel_body(const DGrid& G, uint32_t bid, uint32_t nbid) {
  FOR_ACTIVE(G) {
    uint32_t c = G.slotList[t >> 9] * 512u + (t & 511u);
    G.velG[c * 3] = gvR(G, c * 3);
    G.velG[c * 3 + 1] = gvR(G, c * 3 + 1);
    G.velG[c * 3 + 2] = gvR(G, c * 3 + 2);
  }
}
__global__ void k_save_vel(DGrid G) {
  save_vel_body(G, blockIdx.x, gridDim.x);
}
// nbKill (wgsl-narrowband.js:94-102): calm particles below the band are removed
__global__ void k_nb_kill(const DParams* __restrict__ PP, DGrid G) {
  const DParams& P = *PP;
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  float4 q = G.pos[i];
  if (q.y < ST_DEADY) return;
  // inside the ocean's boundary, or above the ocean surface in its inner half
  const float lw = d_layerW(P, make_float3(q.x, q.y, q.z));
  bool kill = lw >= 0.5f || (lw > 0.25f && q.y > d_oceanEta(P, q.x, q.z) + 0.5f * P.h);
  float4 v = G.vel[i];
  if (!kill && P.nb != 0u && d_calmV(P, make_float3(v.x, v.y, v.z))) {
    int3 cc = d_particleCell(P, make_float3(q.x, q.y, q.z));
    uint32_t sl = G.tmap[d_tileOf(P, cc.x, cc.y, cc.z)];
    if (sl != ST_NONE && d_phiOfBits(G.ctype[d_cellIdx(P, G.tmap, cc.x, cc.y, cc.z)]) < -P.band) kill = true;
  }
  if (kill) {
    G.pos[i] = make_float4(0.0f, -1.0e6f, 0.0f, 0.0f);
    G.vel[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
}
// nbDead (wgsl-narrowband.js:107-110) as a flag per particle id, compacted in id
// order: the reference's particle index is its id (its arrays are never
// reordered by default), and the particle arrays here are in bin order, so the
// k-th dead slot listed is the storage of the k-th dead id
__global__ void k_nb_dead_flags(const DParams* __restrict__ PP, DGrid G, uint32_t* __restrict__ flag) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= PP->n) return;
  flag[G.pid[i]] = G.pos[i].y < ST_DEADY ? 1u : 0u;
}
__global__ void k_nb_dead_list(const DParams* __restrict__ PP, DGrid G, const uint32_t* __restrict__ flag,
                               const uint32_t* __restrict__ start, uint32_t* __restrict__ list) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= PP->n) return;
  const uint32_t id = G.pid[i];
  if (flag[id]) list[start[id]] = i;
}
// nbDeficit (wgsl-narrowband.js:113-130): per active cell (slotList order) how
// many particles it is short; a scan of these gives each cell's entry number and
// the first dead slot it fills, the order the reference's atomics would give
// when its threads run in cell order
__global__ void k_nb_deficit(const DParams* __restrict__ PP, DGrid G, uint32_t* __restrict__ needA, uint32_t* __restrict__ entA,
                             uint32_t ncap) {
  const DParams& P = *PP;
  uint32_t tot = G.ctl[4] * 512u;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < ncap; t += gridDim.x * blockDim.x) {
    uint32_t need = 0u;
    if (t < tot) {
      uint32_t s = G.slotList[t >> 9], local = t & 511u, c = s * 512u + local;
      int3 p = d_slotCoord(P, G.slotTile, s, local);
      uint32_t ct = G.ctype[c];
      float ph = d_phiOfBits(ct);
      bool ok = d_inDom(P, p) && ((ct >> 8) & 0xffu) == 0u;
      if (ok) {
        const float3 cx = d_cellCenter(P, p.x, p.y, p.z);
        const float lw = d_layerW(P, cx);
        if (lw > 0.25f && lw < 0.5f)  // the inner half of the ocean's boundary is always full below the surface
          ok = !(cx.y > d_oceanEta(P, cx.x, cx.z) - 0.25f * P.h || ph < -P.band);
        else
          ok = (ct & 0xffu) == 1u && !(ph > -1.0f) && !(ph < -P.band) && d_calmV(P, d_velGVec(P, G, cx));
      }
      if (ok) {
        uint32_t n = uint32_t(max(G.cnt[c], 0));
        if (!(float(n) >= 0.55f * P.rest)) need = min(st_f2u_sat(rintf(0.85f * P.rest)) - n, 64u);
      }
    }
    needA[t] = need;
    entA[t] = need ? 1u : 0u;
  }
}
__device__ __forceinline__ float d_hash1(uint32_t n0) {
  uint32_t x = n0;
  x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
  return float(x) * 2.3283064e-10f;
}
// nbSpawn (wgsl-narrowband.js:135-149): fill the short cells from the dead slots
__global__ void k_nb_spawn(const DParams* __restrict__ PP, DGrid G, const uint32_t* __restrict__ needA,
                           const uint32_t* __restrict__ needStart, const uint32_t* __restrict__ entIdx,
                           const uint32_t* __restrict__ deadList, const uint32_t* __restrict__ deadCount, uint32_t ncap) {
  const DParams& P = *PP;
  uint32_t tot = G.ctl[4] * 512u, ndead = *deadCount, avail = min(ndead, P.lcap), nent = entIdx[ncap];
  uint32_t maxe = P.lcap / 2u;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += gridDim.x * blockDim.x) {
    uint32_t need = needA[t];
    if (!need) continue;
    uint32_t e = entIdx[t];
    if (e >= maxe) continue;
    uint32_t base = needStart[t];
    // nbReset / nbSpawn's ctl[10]: the needs of the entries the spawn served, summed
    if (e + 1u == min(nent, maxe)) G.ctl[10] = base + need;
    uint32_t s = G.slotList[t >> 9], local = t & 511u;
    int3 p = d_slotCoord(P, G.slotTile, s, local);
    float3 cp = make_float3(float(uint32_t(p.x) & 0xffffu), float(uint32_t(p.y) & 0xffffu), float(uint32_t(p.z) & 0xffffu));
    for (uint32_t m = 0; m < need; m++) {
      uint32_t k = base + m;
      if (k >= avail) break;
      uint32_t i = deadList[k];
      uint32_t hh = (e * 9781u + m * 6271u + 17u) ^ (P.seed * 2654435761u);
      float3 x = f3add(make_float3(P.x0, P.y0, P.z0),
                       f3mul(f3add(cp, make_float3(d_hash1(hh + 1u), d_hash1(hh + 2u), d_hash1(hh + 3u))), P.h));
      float3 v = d_velGVec(P, G, x);
      if (d_layerW(P, x) > 0.25f) v = d_oceanVel(P, x);
      G.pos[i] = make_float4(x.x, x.y, x.z, fminf(f3len(v), 63.0f));
      G.vel[i] = make_float4(v.x, v.y, v.z, 0.0f);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    G.ctl[7] = 0u;
    G.ctl[8] = ndead;
    G.ctl[9] = nent;
    if (nent == 0u) G.ctl[10] = 0u;
  }
}

// ---------------------------------------------------------------------------
// readback in id order (unsort, wgsl-blocks.js:98-99)
__global__ void k_unsort(const float4* __restrict__ src, const uint32_t* __restrict__ pid, float4* __restrict__ out, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[pid[i]] = src[i];
}

// ---------------------------------------------------------------------------
// Batched tanks: one launch per stage for every tank of a batch.  The grid is
// the concatenation of the tanks' own grids; start[t] is the first workgroup of
// tank t (start[nt] the total), so a workgroup finds its tank with one ballot
// over the start table (tanks whose range is empty are passed over) and runs the
// same body as the single-tank kernel with its place within the tank's range.

struct BPos {
  uint32_t t, blk, nblk;
};
__device__ __forceinline__ BPos batch_pos(const uint32_t* __restrict__ start, uint32_t nt) {
  const uint32_t b = blockIdx.x, ln = lane_id();
  uint32_t cnt = 0;
  for (uint32_t base = 0; base < nt; base += 64u) {
    const uint32_t t = base + ln;
    cnt += uint32_t(__popcll(__ballot(t < nt && start[t] <= b)));
    if (cnt < base + 64u) break;  // the last tank starting at or before b is in this chunk
  }
  BPos r;
  r.t = cnt - 1u;
  r.blk = b - start[r.t];
  r.nblk = start[r.t + 1u] - start[r.t];
  return r;
}

// The tank records are read through the constant address space, as kernel
// arguments are: the loads are then invariant, so the compiler reloads a
// pointer where it is used instead of holding it in registers for the whole
// kernel (the single-tank kernels get the same from their by-value DGrid).
typedef const BTank __attribute__((address_space(4))) ConstBTank;
typedef const DParams __attribute__((address_space(4))) ConstDParams;
typedef const DCols __attribute__((address_space(4))) ConstDCols;
#define BATCH_ARGS const BTank* __restrict__ TT, const uint32_t* __restrict__ start, uint32_t nt
#define BATCH_TANK                     \
  const BPos b = batch_pos(start, nt); \
  const BTank& T = *(const BTank*)((ConstBTank*)TT + b.t); \
  const DParams* TP = (const DParams*)(ConstDParams*)T.P;   \
  const DCols* TC = (const DCols*)(ConstDCols*)T.C;         \
  __builtin_assume_dereferenceable(&T, sizeof(BTank));      \
  __builtin_assume_dereferenceable(TP, sizeof(DParams));    \
  __builtin_assume_dereferenceable(TC, sizeof(DCols))
// For g2p the tank's parameter block is copied into LDS once per workgroup:
// read through the pointer above, the compiler cannot treat it as invariant
// there and reads its words with vector loads at each use; from LDS they are
// cheap.  (The other kernels read few words, and their grids can be large.)
__device__ __forceinline__ void batch_stage_params(DParams& dst, const DParams* __restrict__ src) {
  static_assert(sizeof(DParams) % 4 == 0, "word copy");
  for (uint32_t q = threadIdx.x; q < sizeof(DParams) / 4; q += blockDim.x)
    reinterpret_cast<uint32_t*>(&dst)[q] = reinterpret_cast<const uint32_t*>(src)[q];
  __syncthreads();
}

// The cell kernels (one pass per cell of the tank's tiles in use) store into the
// grid between their reads of the parameter block and the grid's pointers.  Read
// through the tank table, neither can be proven unaffected by those stores, so
// each would be loaded again (with vector loads) after every store; the LDS copy
// of the block and a local copy of the pointers are loaded once.
#define BATCH_STAGED            \
  __shared__ DParams sP;        \
  batch_stage_params(sP, TP);   \
  const DGrid G = T.G

// need[0..ntiles) and binCount[0..ntiles] to zero
__global__ void __launch_bounds__(WG) kb_fill_bins(BATCH_ARGS) {
  BATCH_TANK;
  const uint32_t nti = TP->ntiles;
  for (uint32_t i = b.blk * WG + threadIdx.x; i <= nti; i += b.nblk * WG) {
    if (i < nti) T.G.need[i] = 0u;
    T.G.binCount[i] = 0u;
  }
}
__global__ void __launch_bounds__(WG) kb_mark_bin(BATCH_ARGS) {
  BATCH_TANK;
  const DGrid G = T.G;  // local: the body's stores cannot touch it
  mark_bin_body(TP, G, b.blk, b.nblk);
}
__global__ void __launch_bounds__(WG) kb_tile_count(BATCH_ARGS) { BATCH_TANK; tile_count_body(TP, T.G, b.blk, b.nblk); }
__global__ void __launch_bounds__(1024) kb_tile_scan(BATCH_ARGS) {
  BATCH_TANK;
  tile_scan_body(T.G, (TP->ntiles + WG - 1u) / WG, b.blk, b.nblk);
}
__global__ void __launch_bounds__(WG) kb_tile_scatter(BATCH_ARGS) { BATCH_TANK; tile_scatter_body(TP, T.G, b.blk, b.nblk); }
__global__ void kb_tile_alloc(BATCH_ARGS) { BATCH_TANK; tile_alloc_body(T.G, b.blk, b.nblk); }
__global__ void kb_init_new(BATCH_ARGS) { BATCH_TANK; init_new_body(TP, TC, T.G, b.blk, b.nblk); }
__global__ void kb_nbr(BATCH_ARGS) { BATCH_TANK; nbr_body(TP, T.G, b.blk, b.nblk); }
// the bin scan: binCount (ntiles + 1 entries) to binStart
__global__ void __launch_bounds__(WG) kb_scan_blocks(BATCH_ARGS) {
  BATCH_TANK;
  scan_blocks_body(T.G.binCount, T.G.binStart, TP->ntiles + 1u, T.G.scanTmp, b.blk, b.nblk);
}
__global__ void __launch_bounds__(1024) kb_scan_top(BATCH_ARGS) {
  BATCH_TANK;
  const uint32_t n = TP->ntiles + 1u;
  scan_top_body(T.G.scanTmp, (n + WG * 4u - 1u) / (WG * 4u), T.G.binStart, n, b.blk, b.nblk);
}
__global__ void __launch_bounds__(WG) kb_scan_add(BATCH_ARGS) {
  BATCH_TANK;
  scan_add_body(T.G.binStart, TP->ntiles + 1u, T.G.scanTmp, nullptr, b.blk, b.nblk);
}
__global__ void __launch_bounds__(WG) kb_bin_scatter(BATCH_ARGS) {
  BATCH_TANK;
  const DGrid G = T.G;  // local: the body's stores cannot touch it
  bin_scatter_body(TP, G, b.blk, b.nblk);
}
__global__ void kb_clear(BATCH_ARGS) { BATCH_TANK; clear_body(T.G, b.blk, b.nblk); }
__global__ void __launch_bounds__(WG) kb_p2g_tile(BATCH_ARGS) {
  BATCH_TANK;
  const DGrid G = T.G;  // local: the body's stores cannot touch it
  p2g_tile_body(TP, G, b.blk, b.nblk);
}
__global__ void __launch_bounds__(P2G_MERGE_WG) kb_p2g_merge(BATCH_ARGS) {
  BATCH_TANK;
  const DGrid G = T.G;  // local: the body's stores cannot touch it
  p2g_merge_body(G, b.blk, b.nblk);
}
// reduces the partials of the launch whose block table is `of`
__global__ void __launch_bounds__(1024) kb_reduce_stats(BATCH_ARGS, const uint32_t* __restrict__ of, uint32_t which) {
  BATCH_TANK;
  reduce_stats_body(T.G, of[b.t + 1u] - of[b.t], which, b.blk, b.nblk);
}
__global__ void kb_solid_build(BATCH_ARGS) {
  BATCH_TANK;
  const DGrid G = T.G;  // local: the body's stores cannot touch it
  solid_build_body(TP, TC, G, b.blk, b.nblk);
}
// classify keeps reading the parameter block from memory: with the LDS copy it
// took 104 VGPRs and ran slower (77 against 68 us per 64-tank substep)
__global__ void __launch_bounds__(WG) kb_classify(BATCH_ARGS) {
  BATCH_TANK;
  const DGrid G = T.G;  // local: the body's stores cannot touch it
  classify_body(TP, G, b.blk, b.nblk);
}
__global__ void kb_norm_faces(BATCH_ARGS) { BATCH_TANK; BATCH_STAGED; norm_faces_body(&sP, TC, G, b.blk, b.nblk); }
__global__ void kb_divergence(BATCH_ARGS) { BATCH_TANK; BATCH_STAGED; divergence_body(&sP, G, b.blk, b.nblk); }
// one SOR iteration, pres -> pres2 (odd = 0) or back; every tank's range is a
// multiple of 8 workgroups starting at a multiple of 8, so a workgroup's place
// within its tank keeps the XCD it runs on
__global__ void __launch_bounds__(WG) kb_sor_pipe(BATCH_ARGS, uint32_t odd) {
  BATCH_TANK;
  sor_pipe_body(TP, T.G, odd ? T.G.pres2 : T.G.pres, odd ? T.G.pres : T.G.pres2, ST_NONE, b.blk, b.nblk);
}
__global__ void kb_copy_tiles(BATCH_ARGS) { BATCH_TANK; copy_tiles_body(T.G, T.G.pres2, T.G.pres, ST_NONE, b.blk, b.nblk); }
// the path word of the solve just done: one launch per iteration
__global__ void kb_sor_path(BATCH_ARGS) {
  BATCH_TANK;
  if (threadIdx.x == 0 && b.blk == 0) T.G.sorBar[COOP_MAX_WG + 3] = 0u;
}
// Batched resident SOR.  The tanks' solves run in one cooperative launch of
// the resident workgroups (one per CU): each tank gets a range of them, in one
// of a few phases when all tanks' tiles do not fit in LDS at once, and runs
// sor_resident_body over its range (its own halo entries, tags and barrier
// words).  A tank's workgroups only wait on each other, and every workgroup
// goes through the phases in order, so a phase never waits on a later one.
// Per tank active this substep (its range in the first iteration's table is
// not empty): kb_coop_lists (one 1024-thread workgroup per tank) builds its
// Morton-ordered slot list, as k_coop_flags / scan / k_coop_compact do, and
// clears its barrier words; kb_sor_plan (one thread) assigns the workgroups.
__global__ void __launch_bounds__(1024) kb_coop_lists(const BTank* __restrict__ TT, const uint32_t* __restrict__ act, uint32_t nt) {
  const uint32_t t = blockIdx.x, tid = threadIdx.x;
  if (t >= nt || act[t + 1] == act[t]) return;
  const DGrid G = TT[t].G;
  const uint32_t ntl = TT[t].P->ntiles;
  __shared__ uint32_t wsum[1024 / 64];
  for (uint32_t q = tid; q < COOP_MAX_WG; q += 1024u) G.sorBar[q] = 0u;
  if (tid == 0u) G.sorBar[COOP_MAX_WG + 2] = 0u;
  uint32_t base = 0;
  for (uint32_t m0 = 0; m0 < ntl; m0 += 1024u) {
    const uint32_t m = m0 + tid;
    const uint32_t sl = m < ntl ? G.tmap[G.mortonTiles[m]] : ST_NONE;
    const uint32_t f = sl != ST_NONE ? 1u : 0u;
    const uint64_t bal = __ballot(f != 0u);
    const uint32_t inWave = uint32_t(__popcll(bal & lanemask_lt()));
    if (lane_id() == 0u) wsum[tid >> 6] = uint32_t(__popcll(bal));
    __syncthreads();
    uint32_t before = 0, tot = 0;
    for (uint32_t u = 0; u < 1024u / 64u; u++) {
      if (u < (tid >> 6)) before += wsum[u];
      tot += wsum[u];
    }
    if (f) G.coopList[base + before + inWave] = sl;
    base += tot;
    __syncthreads();
  }
}

// plan[3 t .. 3 t + 2] = (phase, first workgroup, workgroups) of tank t (phase
// ST_NONE: no solve), plan[3 nt] = phases.  First fit in tank order, each tank
// at least the workgroups its tiles need in LDS (all `grid` when even those
// are too few: it then streams); a phase's spare workgroups go to its tanks in
// proportion to their tiles.  A tank that streams is alone in its phase with
// all `grid` workgroups (at least 2, which its tile pipelines need).
__global__ void kb_sor_plan(const BTank* __restrict__ TT, const uint32_t* __restrict__ act, uint32_t nt, uint32_t grid,
                            uint32_t* __restrict__ plan) {
  if (threadIdx.x != 0u || blockIdx.x != 0u) return;
  uint32_t phase = 0, used = 0, first = 0;
  uint64_t tiles = 0;
  auto close = [&](uint32_t endT) {  // spread the spare workgroups of the phase over tanks [first, endT)
    const uint32_t spare = grid - used;
    uint32_t at = 0;
    for (uint32_t q = first; q < endT; q++) {
      if (plan[3 * q] != phase) continue;
      const uint32_t nact = TT[q].G.ctl[4];
      const uint32_t extra = tiles ? uint32_t(uint64_t(spare) * nact / tiles) : 0u;
      plan[3 * q + 1] = at;
      plan[3 * q + 2] += extra;
      at += plan[3 * q + 2];
    }
  };
  for (uint32_t t = 0; t < nt; t++) {
    plan[3 * t] = ST_NONE;
    plan[3 * t + 1] = 0u;
    plan[3 * t + 2] = 0u;
    const uint32_t nact = act[t + 1] != act[t] ? TT[t].G.ctl[4] : 0u;
    if (nact == 0u) continue;
    const uint32_t need = min(grid, (nact + COOP_TMAX - 1u) / COOP_TMAX);
    if (used + need > grid) {
      close(t);
      phase++;
      used = 0;
      tiles = 0;
      first = t;
    }
    plan[3 * t] = phase;
    plan[3 * t + 2] = need;
    used += need;
    tiles += nact;
  }
  if (used) {
    close(nt);
    phase++;
  }
  plan[3 * nt] = phase;
}

__global__ void __launch_bounds__(COOP_WG) kb_sor_resident(const BTank* __restrict__ TT, const uint32_t* __restrict__ plan, uint32_t nt) {
  const uint32_t nph = plan[3 * nt], w = blockIdx.x;
  for (uint32_t p = 0; p < nph; p++) {
    uint32_t t = ST_NONE;
    for (uint32_t q = 0; q < nt; q++)
      if (plan[3 * q] == p && w >= plan[3 * q + 1] && w < plan[3 * q + 1] + plan[3 * q + 2]) t = q;
    if (t == ST_NONE) continue;
    const BTank& B = TT[t];
    sor_resident_body(B.P, B.G, B.P->iters, 0u, 0u, w - plan[3 * t + 1], plan[3 * t + 2]);
    __syncthreads();  // the LDS of this phase's tank is free again
  }
}

__global__ void kb_project(BATCH_ARGS) { BATCH_TANK; BATCH_STAGED; project_body(&sP, G, b.blk, b.nblk); }
__global__ void kb_foam_table(BATCH_ARGS) { BATCH_TANK; foam_table_body(TP, T.G, b.blk, b.nblk); }
__global__ void __launch_bounds__(WG, 4) kb_g2p_tile(BATCH_ARGS) {
  BATCH_TANK;
  __shared__ DParams sP;
  batch_stage_params(sP, TP);
  const DGrid G = T.G;  // local: the body's stores cannot touch it
  g2p_tile_body(&sP, TC, G, 1u, b.blk, b.nblk);
}

// ---------------------------------------------------------------------------
// launch wrappers

static inline uint32_t gridFor(size_t items, uint32_t block) { return uint32_t((items + block - 1) / block); }

void launch_fill_u32(hipStream_t st, uint32_t* p, size_t n, uint32_t v) {
  if (!n) return;
  uint32_t g = std::min<uint32_t>(gridFor(n, WG), 8192u);
  hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(WG), 0, st, p, n, v);
}

static void launch_scan(hipStream_t st, const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* sums, uint32_t* copy) {
  uint32_t nb = gridFor(n, WG * 4);
  hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(WG), 0, st, in, out, n, sums);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, sums, nb, out, n);
  hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(WG), 0, st, out, n, (const uint32_t*)sums, copy);
}

void launch_blocks(hipStream_t st, const LaunchCtx& L) {
  uint32_t nt = L.ntiles;
  if (!L.binned) {  // else the last g2p already marked and counted A's positions
    launch_fill_u32(st, L.G.need, nt, 0u);
    launch_fill_u32(st, L.G.binCount, size_t(nt) + 1, 0u);
    if (L.n) hipLaunchKernelGGL(k_mark_bin, dim3(gridFor(L.n, WG)), dim3(WG), 0, st, L.P, L.G);
  }
  // narrow band: last step's tiles holding water by depth, and on a start the
  // tiles the starting water reaches (both before the counters reset)
  if (L.nbOn && L.initStep != 1) hipLaunchKernelGGL(k_nb_mark_phi, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.G);
  if (L.initStep) hipLaunchKernelGGL(k_nb_mark_fill, dim3(gridFor(nt, WG)), dim3(WG), 0, st, L.P, L.C, L.G);
  uint32_t nwg = gridFor(nt, WG);
  hipLaunchKernelGGL(k_tile_count, dim3(nwg), dim3(WG), 0, st, L.P, L.G);
  hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, st, L.G, nwg);
  hipLaunchKernelGGL(k_tile_scatter, dim3(nwg), dim3(WG), 0, st, L.P, L.G);
  hipLaunchKernelGGL(k_tile_alloc, dim3(std::max<uint32_t>(1u, std::min<uint32_t>(gridFor(nt, WG), 256u))), dim3(WG), 0, st, L.G);
  hipLaunchKernelGGL(k_init_new, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.C, L.G);
  hipLaunchKernelGGL(k_nbr, dim3(std::max<uint32_t>(1u, std::min<uint32_t>(gridFor(size_t(L.scap) * 27, WG), 2048u))), dim3(WG), 0, st, L.P, L.G);
  if (L.initStep) hipLaunchKernelGGL(k_nb_phi_init, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.C, L.G);
  // particle bins: exclusive scan of the counts, then the bin list
  launch_scan(st, L.G.binCount, L.G.binStart, nt + 1, L.G.scanTmp, nullptr);
  if (L.n) hipLaunchKernelGGL(k_bin_scatter, dim3(gridFor(L.n, WG)), dim3(WG), 0, st, L.P, L.G);
}

void launch_clear(hipStream_t st, const LaunchCtx& L) {
  hipLaunchKernelGGL(k_clear, dim3(512 / WG), dim3(WG), 0, st, L.G);
}

void launch_p2g(hipStream_t st, const LaunchCtx& L) {
  if (!L.n) return;
  hipLaunchKernelGGL(k_p2g_tile, dim3(L.tileGrid), dim3(WG), 0, st, L.P, L.G);
  hipLaunchKernelGGL(k_p2g_merge, dim3(L.tileGrid), dim3(P2G_MERGE_WG), 0, st, L.G);
  hipLaunchKernelGGL(k_reduce_stats, dim3(1), dim3(1024), 0, st, L.G, L.tileGrid, 0u);
}

void launch_classify(hipStream_t st, const LaunchCtx& L, bool solid) {
  if (solid) hipLaunchKernelGGL(k_solid_build, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.C, L.G);
  hipLaunchKernelGGL(k_classify, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.G);
  hipLaunchKernelGGL(k_reduce_stats, dim3(1), dim3(1024), 0, st, L.G, L.cellGrid, 1u);
  if (L.nbOn) hipLaunchKernelGGL(k_nb_vel, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.G);
}

void launch_narrowband(hipStream_t st, const LaunchCtx& L) {
  hipLaunchKernelGGL(k_nb_phi, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.G);
  hipLaunchKernelGGL(k_nb_jac, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.G, 1u);
  hipLaunchKernelGGL(k_nb_jac, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.G, 0u);
  hipLaunchKernelGGL(k_nb_jac, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.G, 1u);
}

void launch_reseed(hipStream_t st, const LaunchCtx& L, const NbScratch& S) {
  if (!L.n) return;
  uint32_t gp = gridFor(L.n, WG);
  hipLaunchKernelGGL(k_nb_kill, dim3(gp), dim3(WG), 0, st, L.P, L.G);
  hipLaunchKernelGGL(k_nb_dead_flags, dim3(gp), dim3(WG), 0, st, L.P, L.G, S.deadFlag);
  launch_scan(st, S.deadFlag, S.deadStart, L.n, S.sums, nullptr);
  hipLaunchKernelGGL(k_nb_dead_list, dim3(gp), dim3(WG), 0, st, L.P, L.G, (const uint32_t*)S.deadFlag, (const uint32_t*)S.deadStart,
                     S.deadList);
  hipLaunchKernelGGL(k_nb_deficit, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.G, S.need, S.ent, S.ncap);
  launch_scan(st, S.need, S.needStart, S.ncap, S.sums, nullptr);
  launch_scan(st, S.ent, S.entIdx, S.ncap, S.sums, nullptr);
  hipLaunchKernelGGL(k_nb_spawn, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.G, (const uint32_t*)S.need, (const uint32_t*)S.needStart,
                     (const uint32_t*)S.entIdx, (const uint32_t*)S.deadList, (const uint32_t*)(S.deadStart + L.n), S.ncap);
}

void launch_norm_faces(hipStream_t st, const LaunchCtx& L) {
  hipLaunchKernelGGL(k_norm_faces, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.C, L.G);
}

void launch_divergence(hipStream_t st, const LaunchCtx& L) {
  hipLaunchKernelGGL(k_divergence, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.G);
}

uint32_t sor_resident_grid(uint32_t cus) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sor_pipe, WG, 0) != hipSuccess || nb < 1) nb = 1;
  uint32_t g = (cus * uint32_t(nb)) & ~7u;
  return g < 8u ? 8u : g;
}

uint32_t sor_coop_grid(uint32_t cus) {
  int dev = 0, coop = 0, nb = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess || !coop)
    return 0u;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sor_resident, COOP_WG, 0) != hipSuccess || nb < 1) return 0u;
  return std::min<uint32_t>(cus, COOP_MAX_WG);
}

void morton_tiles(uint32_t tx, uint32_t ty, uint32_t tz, uint32_t* out) {
  auto spread = [](uint64_t v) {  // bit i of v to bit 3i
    uint64_t r = 0;
    for (int b = 0; b < 21; b++) r |= ((v >> b) & 1ull) << (3 * b);
    return r;
  };
  std::vector<std::pair<uint64_t, uint32_t>> key;
  key.reserve(size_t(tx) * ty * tz);
  for (uint32_t k = 0; k < tz; k++)
    for (uint32_t j = 0; j < ty; j++)
      for (uint32_t i = 0; i < tx; i++)
        key.emplace_back(spread(i) | (spread(j) << 1) | (spread(k) << 2), i + tx * (j + ty * k));
  std::sort(key.begin(), key.end());
  for (size_t m = 0; m < key.size(); m++) out[m] = key[m].second;
}

hipError_t launch_pressure(hipStream_t st, const LaunchCtx& L, int iters) {
  hipError_t coopErr = hipSuccess;
  if (L.coopGrid && iters > 0) {
    // active slots in Morton order, cleared barrier flags, then the whole solve
    // in one cooperative launch (resident when the tiles fit in LDS, streaming
    // tile windows otherwise)
    uint32_t nt = L.ntiles, g = gridFor(nt, WG);
    hipLaunchKernelGGL(k_coop_flags, dim3(g), dim3(WG), 0, st, L.P, L.G);
    launch_scan(st, L.G.coopFlags, L.G.coopOffs, nt, L.G.scanTmp, nullptr);
    hipLaunchKernelGGL(k_coop_compact, dim3(g), dim3(WG), 0, st, L.P, L.G);
    launch_fill_u32(st, L.G.sorBar, COOP_MAX_WG, 0u);
    launch_fill_u32(st, L.G.sorBar + COOP_MAX_WG + 2, 1, 0u);
    const DParams* P = L.P;
    DGrid G = L.G;
    uint32_t it = uint32_t(iters), seq = L.sorSeq, stream = L.sorStream ? 1u : 0u;
    void* args[] = {(void*)&P, (void*)&G, (void*)&it, (void*)&seq, (void*)&stream};
    coopErr = hipLaunchCooperativeKernel((const void*)k_sor_resident, dim3(L.coopGrid), dim3(COOP_WG), args, 0, st);
    if (coopErr == hipSuccess) return hipSuccess;
    (void)hipGetLastError();  // the per-iteration launches below do the solve
  }
  // one iteration per launch, ping-pong between the two pressure buffers; an
  // odd count leaves the result in pres2 and it is copied back.
  for (int it = 0; it < iters; it++) {
    const float* pin = (it & 1) ? L.G.pres2 : L.G.pres;
    float* pout = (it & 1) ? L.G.pres : L.G.pres2;
    hipLaunchKernelGGL(k_sor_pipe, dim3(L.pipeGrid), dim3(WG), 0, st, L.P, L.G, pin, pout, L.sorSeq);
  }
  if (iters & 1)
    hipLaunchKernelGGL(k_copy_tiles, dim3(L.cellGrid), dim3(WG), 0, st, L.G, (const float*)L.G.pres2, L.G.pres, L.sorSeq);
  launch_fill_u32(st, L.G.sorBar + COOP_MAX_WG + 3, 1, 0u);
  return coopErr;
}

void launch_project(hipStream_t st, const LaunchCtx& L) {
  hipLaunchKernelGGL(k_project, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.G);
  if (L.extrap && L.exMask) {
    const size_t C = size_t(L.scap) * 512u;
    hipLaunchKernelGGL(k_extrap_init, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.G, L.exMask);
    for (uint32_t l = 0; l < L.extrap; l++)
      hipLaunchKernelGGL(k_extrap_layer, dim3(L.cellGrid), dim3(WG), 0, st, L.P, L.G, (const uint8_t*)(L.exMask + (l & 1u) * C),
                         L.exMask + ((l + 1u) & 1u) * C);
  }
  if (L.nbOn) hipLaunchKernelGGL(k_save_vel, dim3(L.cellGrid), dim3(WG), 0, st, L.G);
}

void launch_g2p(hipStream_t st, const LaunchCtx& L) {
  if (!L.n) return;
  if (L.binNext) {  // the next substep's tile marks and bin counts start here
    launch_fill_u32(st, L.G.need, L.ntiles, 0u);
    launch_fill_u32(st, L.G.binCount, size_t(L.ntiles) + 1, 0u);
  }
  hipLaunchKernelGGL(k_foam_table, dim3(FOAM_TAB / WG), dim3(WG), 0, st, L.P, L.G);
  uint32_t g = std::min<uint32_t>(L.ntiles + 1, 8192u);
  hipLaunchKernelGGL(k_g2p_tile, dim3(g), dim3(WG), 0, st, L.P, L.C, L.G, L.binNext ? 1u : 0u);
}

void launch_unsort(hipStream_t st, const float4* src, const uint32_t* pid, float4* out, uint32_t n) {
  if (!n) return;
  hipLaunchKernelGGL(k_unsort, dim3(gridFor(n, WG)), dim3(WG), 0, st, src, pid, out, n);
}

// ---------------------------------------------------------------------------
// batched substep (host side)

size_t batch_table_words(uint32_t nt, uint32_t itersMax) { return size_t(BT_PIPE0 + itersMax) * (nt + 1u); }

// Grid-stride kernels take any number of workgroups per tank.  A single tank
// launches up to 2 per slot (cells), 1 per slot (p2g) or 1 per bin (g2p); across
// a batch that many would mostly find no tile in use and exit, so the batch
// splits a fixed total over the tanks by pool size (at least one each, never
// more than the single-tank grid).
#define BATCH_CELL_WG 4096u
#define BATCH_TILE_WG 16384u
static uint32_t batch_share(uint64_t total, uint64_t part, uint64_t whole, uint32_t most) {
  uint64_t c = whole ? (total * part + whole - 1u) / whole : 1u;
  return uint32_t(std::max<uint64_t>(1u, std::min<uint64_t>(c, most)));
}

void batch_tables(const BTankPlan* plan, uint32_t nt, uint32_t itersMax, uint32_t sorResident, uint32_t* tab) {
  uint64_t scapTot = 0, binTot = 0;
  for (uint32_t t = 0; t < nt; t++)
    if (plan[t].n) {
      scapTot += plan[t].scap;
      binTot += plan[t].ntiles + 1u;
    }
  const size_t w = size_t(nt) + 1u;
  for (uint32_t k = 0; k < uint32_t(BT_PIPE0) + itersMax; k++) {
    uint32_t* s = tab + k * w;
    uint32_t acc = 0;
    for (uint32_t t = 0; t < nt; t++) {
      const BTankPlan& p = plan[t];
      uint32_t c = 0;
      if (p.n) {
        const uint32_t cell = batch_share(BATCH_CELL_WG, p.scap, scapTot, std::max<uint32_t>(1u, std::min<uint32_t>(p.scap * 2u, 4096u)));
        const uint32_t nwg = gridFor(p.ntiles, WG), bins = gridFor(size_t(p.ntiles) + 1u, WG);
        switch (k) {
          case BT_FILL_MARK: c = p.binned ? 0u : bins; break;
          case BT_MARK: c = p.binned ? 0u : gridFor(p.n, WG); break;
          case BT_NWG: c = nwg; break;
          case BT_ONE: c = 1u; break;
          case BT_ALLOC: c = std::max<uint32_t>(1u, std::min<uint32_t>(nwg, 256u)); break;
          case BT_CELL: c = cell; break;
          case BT_NBR: c = std::max<uint32_t>(1u, std::min<uint32_t>(gridFor(size_t(p.scap) * 27u, WG), 2048u)); break;
          case BT_SCAN: c = gridFor(size_t(p.ntiles) + 1u, WG * 4u); break;
          case BT_PART: c = gridFor(p.n, WG); break;
          case BT_TILE: c = batch_share(BATCH_TILE_WG, p.scap, scapTot, std::max<uint32_t>(1u, std::min<uint32_t>(p.scap, 8192u))); break;
          case BT_ONE_N: c = 1u; break;
          case BT_SOLID: c = p.solid ? cell : 0u; break;
          case BT_COPY: c = (p.iters & 1u) ? cell : 0u; break;
          case BT_FOAM: c = FOAM_TAB / WG; break;
          case BT_G2P: c = batch_share(BATCH_TILE_WG, p.ntiles + 1u, binTot, std::min<uint32_t>(p.ntiles + 1u, 8192u)); break;
          case BT_FILL_NEXT: c = bins; break;
          default: {
            // SOR: the tank's share of the resident workgroups by pool size,
            // whole groups of 8, at least one group and at most one workgroup
            // per slot (rounded up to 8)
            if (k - uint32_t(BT_PIPE0) >= p.iters) break;
            const uint32_t most = (p.scap + 7u) & ~7u;
            uint32_t share = uint32_t((uint64_t(sorResident) * p.scap + scapTot - 1u) / scapTot);
            share = (share + 7u) & ~7u;
            c = std::max<uint32_t>(8u, std::min<uint32_t>(share, most));
          }
        }
      }
      s[t] = acc;
      acc += c;
    }
    s[nt] = acc;
  }
}

bool launch_batch_step(hipStream_t st, const BTank* dT, const uint32_t* dtab, const uint32_t* htab, uint32_t nt, uint32_t itersMax,
                       uint32_t coopGrid, uint32_t* dPlan) {
  const size_t w = size_t(nt) + 1u;
  auto tab = [&](uint32_t k) { return dtab + k * w; };
  auto tot = [&](uint32_t k) { return htab[k * w + nt]; };
#define BLAUNCH(kern, k, threads, ...) \
  if (tot(k)) hipLaunchKernelGGL(kern, dim3(tot(k)), dim3(threads), 0, st, dT, tab(k), nt, ##__VA_ARGS__)
  // blocks: marks and bins (unless the last g2p made them), tile allocation,
  // neighbour slots, bin scan and bin list
  BLAUNCH(kb_fill_bins, BT_FILL_MARK, WG);
  BLAUNCH(kb_mark_bin, BT_MARK, WG);
  BLAUNCH(kb_tile_count, BT_NWG, WG);
  BLAUNCH(kb_tile_scan, BT_ONE, 1024);
  BLAUNCH(kb_tile_scatter, BT_NWG, WG);
  BLAUNCH(kb_tile_alloc, BT_ALLOC, WG);
  BLAUNCH(kb_init_new, BT_CELL, WG);
  BLAUNCH(kb_nbr, BT_NBR, WG);
  BLAUNCH(kb_scan_blocks, BT_SCAN, WG);
  BLAUNCH(kb_scan_top, BT_ONE, 1024);
  BLAUNCH(kb_scan_add, BT_SCAN, WG);
  BLAUNCH(kb_bin_scatter, BT_PART, WG);
  // clear, p2g
  BLAUNCH(kb_clear, BT_CELL, WG);
  BLAUNCH(kb_p2g_tile, BT_TILE, WG);
  BLAUNCH(kb_p2g_merge, BT_TILE, P2G_MERGE_WG);
  BLAUNCH(kb_reduce_stats, BT_ONE_N, 1024, tab(BT_TILE), 0u);
  // classify, normFaces, divergence
  BLAUNCH(kb_solid_build, BT_SOLID, WG);
  BLAUNCH(kb_classify, BT_CELL, WG);
  BLAUNCH(kb_reduce_stats, BT_ONE, 1024, tab(BT_CELL), 1u);
  BLAUNCH(kb_norm_faces, BT_CELL, WG);
  BLAUNCH(kb_divergence, BT_CELL, WG);
  // pressure: the resident solves of all tanks in one cooperative launch, or
  // one launch per iteration (ping-pong between pres and pres2)
  bool resident = false;
  if (coopGrid && dPlan && itersMax) {
    hipLaunchKernelGGL(kb_coop_lists, dim3(nt), dim3(1024), 0, st, dT, tab(BT_PIPE0), nt);
    hipLaunchKernelGGL(kb_sor_plan, dim3(1), dim3(64), 0, st, dT, tab(BT_PIPE0), nt, coopGrid, dPlan);
    const BTank* a0 = dT;
    const uint32_t* a1 = dPlan;
    uint32_t a2 = nt;
    void* args[] = {(void*)&a0, (void*)&a1, (void*)&a2};
    resident = hipLaunchCooperativeKernel((const void*)kb_sor_resident, dim3(coopGrid), dim3(COOP_WG), args, 0, st) == hipSuccess;
    if (!resident) (void)hipGetLastError();  // the launches per iteration below do the solve
  }
  if (!resident) {
    for (uint32_t it = 0; it < itersMax; it++) BLAUNCH(kb_sor_pipe, uint32_t(BT_PIPE0) + it, WG, it & 1u);
    BLAUNCH(kb_copy_tiles, BT_COPY, WG);
    BLAUNCH(kb_sor_path, BT_ONE, 64);
  }
  // project, g2p (which marks and bins the next substep)
  BLAUNCH(kb_project, BT_CELL, WG);
  BLAUNCH(kb_fill_bins, BT_FILL_NEXT, WG);
  BLAUNCH(kb_foam_table, BT_FOAM, WG);
  BLAUNCH(kb_g2p_tile, BT_G2P, WG);
#undef BLAUNCH
  return resident;
}

// ---------------------------------------------------------------------------
// Whitewater and the air field (src/flip/wgsl-whitewater.js:1-163): one pool of
// smax points that never push the water -- spray, mist, foam and bubbles --
// thrown off by fast exposed water and churn, moved by kind each substep; and a
// coarse air field the spray drags toward.  Slots are handed out as a ring
// (sctl[0] counts them); until the ring wraps only the first sctl[0] can be
// alive, so the point kernels loop over min(sctl[0], smax) slots read on the
// device.  New points take consecutive slots in particle-id order (an exclusive
// scan of each particle's count), where the reference takes them in the order
// its atomics land; a slot that still holds a living point is skipped, as there.

#define WWHOLD 0.2f

// wetCell / wetAt (wgsl-whitewater.js:7-22): how much water is at a cell, 0..1
__device__ __forceinline__ float d_wetCell(const DParams& P, const DGrid& G, int32_t i, int32_t j, int32_t k) {
  if (!d_inDom(P, make_int3(i, j, k))) return 0.0f;
  uint32_t s = G.tmap[d_tileOf(P, i, j, k)];
  if (s == ST_NONE) return 0.0f;
  uint32_t c = d_cellIdx(P, G.tmap, i, j, k);
  if (P.nb != 0u && (G.ctype[c] & 0xffu) == 1u && d_phiOfBits(G.ctype[c]) < -1.0f) return 1.0f;
  return fminf(float(G.cnt[c]) / P.rest, 1.0f);
}
__device__ __forceinline__ float d_wetAt(const DParams& P, const DGrid& G, float3 p) {
  float3 g = make_float3((p.x - P.x0) * P.ih - 0.5f, (p.y - P.y0) * P.ih - 0.5f, (p.z - P.z0) * P.ih - 0.5f);
  float3 f0 = make_float3(floorf(g.x), floorf(g.y), floorf(g.z));
  float3 fr = f3sub(g, f0);
  int32_t i0 = st_f2i_sat(f0.x), j0 = st_f2i_sat(f0.y), k0 = st_f2i_sat(f0.z);
  float s = 0.0f;
  for (uint32_t c = 0; c < 8u; c++) {
    int32_t dx = int32_t(c & 1u), dy = int32_t((c >> 1) & 1u), dz = int32_t((c >> 2) & 1u);
    float w = (dx == 1 ? fr.x : 1.0f - fr.x) * (dy == 1 ? fr.y : 1.0f - fr.y) * (dz == 1 ? fr.z : 1.0f - fr.z);
    if (w > 0.0f) s = s + w * d_wetCell(P, G, i0 + dx, j0 + dy, k0 + dz);
  }
  return s;
}
// air grid (wgsl-whitewater.js:23-50): cells of airH from the box's low corner
__device__ __forceinline__ uint32_t d_airIdx(const DParams& P, int32_t i, int32_t j, int32_t k) {
  i = d_iclamp(i, 0, int32_t(P.aNx) - 1);
  j = d_iclamp(j, 0, int32_t(P.aNy) - 1);
  k = d_iclamp(k, 0, int32_t(P.aNz) - 1);
  return uint32_t(i) + P.aNx * (uint32_t(j) + P.aNy * uint32_t(k));
}
__device__ __forceinline__ float3 d_airAt(const DParams& P, const float4* __restrict__ air, float3 p) {
  float3 g = make_float3((p.x - P.x0) / P.airH - 0.5f, (p.y - P.y0) / P.airH - 0.5f, (p.z - P.z0) / P.airH - 0.5f);
  float3 f0 = make_float3(floorf(g.x), floorf(g.y), floorf(g.z));
  float3 fr = f3sub(g, f0);
  int32_t i0 = st_f2i_sat(f0.x), j0 = st_f2i_sat(f0.y), k0 = st_f2i_sat(f0.z);
  float3 s = make_float3(0.0f, 0.0f, 0.0f);
  for (uint32_t c = 0; c < 8u; c++) {
    int32_t dx = int32_t(c & 1u), dy = int32_t((c >> 1) & 1u), dz = int32_t((c >> 2) & 1u);
    float w = (dx == 1 ? fr.x : 1.0f - fr.x) * (dy == 1 ? fr.y : 1.0f - fr.y) * (dz == 1 ? fr.z : 1.0f - fr.z);
    s = f3add(s, f3mul(c3(air[d_airIdx(P, i0 + dx, j0 + dy, k0 + dz)]), w));
  }
  return s;
}
__device__ __forceinline__ void d_airPush(const DParams& P, const DGrid& G, float3 p, float3 v, float w) {
  int32_t ci = st_f2i_sat(floorf((p.x - P.x0) / P.airH)), cj = st_f2i_sat(floorf((p.y - P.y0) / P.airH)),
          ck = st_f2i_sat(floorf((p.z - P.z0) / P.airH));
  uint32_t c = d_airIdx(P, ci, cj, ck) * 4u;
  atomicAdd(&G.airAcc[c], st_f2i_sat(v.x * w * 256.0f));
  atomicAdd(&G.airAcc[c + 1u], st_f2i_sat(v.y * w * 256.0f));
  atomicAdd(&G.airAcc[c + 2u], st_f2i_sat(v.z * w * 256.0f));
  atomicAdd(&G.airAcc[c + 3u], st_f2i_sat(w * 256.0f));
}
__device__ __forceinline__ uint32_t d_wwUsed(const DParams& P, const DGrid& G) { return min(G.sctl[0], P.smax); }

// airSplatW (54-59): water near the surface drags the air
__global__ void k_air_splat_w(const DParams* __restrict__ PP, DGrid G) {
  const DParams& P = *PP;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < P.n; i += gridDim.x * blockDim.x) {
    float4 q = G.pos[i];
    if (q.y < ST_DEADY) continue;
    float3 p = c3(q);
    int3 c = d_particleCell(P, p);
    if (G.tmap[d_tileOf(P, c.x, c.y, c.z)] != ST_NONE && float(G.cnt[d_cellIdx(P, G.tmap, c.x, c.y, c.z)]) > 0.85f * P.rest) continue;
    d_airPush(P, G, p, c3(G.vel[i]), 1.0f);
  }
}
// airSplatS (61-66): flying spray drags it too
__global__ void k_air_splat_s(const DParams* __restrict__ PP, DGrid G) {
  const DParams& P = *PP;
  const uint32_t used = d_wwUsed(P, G);
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < used; j += gridDim.x * blockDim.x) {
    float4 sp = G.sPos[j];
    if (sp.w <= 0.0f || G.sVel[j].w > 1.5f) continue;
    d_airPush(P, G, c3(sp), c3(G.sVel[j]), 0.25f);
  }
}
// airUpdate (68-86): spread, pulled toward what pushed it, moved by objects, settled to the wind
__global__ void k_air_update(const DParams* __restrict__ PP, const DCols* __restrict__ C, DGrid G) {
  const DParams& P = *PP;
  const uint32_t n = P.aNx * P.aNy * P.aNz;
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    int32_t i = int32_t(c % P.aNx), j = int32_t((c / P.aNx) % P.aNy), k = int32_t(c / (P.aNx * P.aNy));
    float3 a = c3(G.air[c]);
    float3 nb = f3add(f3add(f3add(f3add(f3add(c3(G.air[d_airIdx(P, i + 1, j, k)]), c3(G.air[d_airIdx(P, i - 1, j, k)])),
                                            c3(G.air[d_airIdx(P, i, j + 1, k)])),
                                      c3(G.air[d_airIdx(P, i, j - 1, k)])),
                                c3(G.air[d_airIdx(P, i, j, k + 1)])),
                          c3(G.air[d_airIdx(P, i, j, k - 1)]));
    nb = f3div(nb, 6.0f);
    a = make_float3(st_mixf(a.x, nb.x, P.airBlur), st_mixf(a.y, nb.y, P.airBlur), st_mixf(a.z, nb.z, P.airBlur));
    float w = float(G.airAcc[c * 4u + 3u]) / 256.0f;
    if (w > 0.0f) {
      float3 vt = f3div(make_float3(float(G.airAcc[c * 4u]), float(G.airAcc[c * 4u + 1u]), float(G.airAcc[c * 4u + 2u])), w * 256.0f);
      float t = 1.0f - st_expf(-P.airCouple * fminf(w, 8.0f) * P.dt);
      a = make_float3(st_mixf(a.x, vt.x, t), st_mixf(a.y, vt.y, t), st_mixf(a.z, vt.z, t));
    }
    float3 x = make_float3(P.x0 + (float(i) + 0.5f) * P.airH, P.y0 + (float(j) + 0.5f) * P.airH, P.z0 + (float(k) + 0.5f) * P.airH);
    for (uint32_t q = 0; q < P.ncols; q++) {
      if (!d_colNear(P, C, q, x)) continue;
      float phi;
      float3 wn;
      d_sdfEval(C, G.sdf, q, x, phi, wn);
      if (phi < 0.6f * P.airH) {
        a = d_colSurfVel(C, q, x);
        break;
      }
    }
    float t = 1.0f - st_expf(-P.airDecay * P.dt);
    a = make_float3(st_mixf(a.x, P.q1, t), st_mixf(a.y, P.q2, t), st_mixf(a.z, P.q3, t));
    G.air2[c] = make_float4(a.x, a.y, a.z, 0.0f);
  }
}
__global__ void k_air_copy(const DParams* __restrict__ PP, DGrid G) {
  const DParams& P = *PP;
  const uint32_t n = P.aNx * P.aNy * P.aNz;
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) G.air[c] = G.air2[c];
}
// wwSample (88-95): how wet each point is and the water's velocity there
__global__ void k_ww_sample(const DParams* __restrict__ PP, DGrid G) {
  const DParams& P = *PP;
  const uint32_t used = d_wwUsed(P, G);
  GlobalGrid gg{P, G};
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < used; j += gridDim.x * blockDim.x) {
    float4 sp = G.sPos[j];
    if (sp.w <= 0.0f) continue;
    float3 p = c3(sp), vf = make_float3(0.0f, 0.0f, 0.0f);
    int3 pc = d_particleCell(P, p);
    if (G.tmap[d_tileOf(P, pc.x, pc.y, pc.z)] != ST_NONE) {
      float3 vo;
      gg.sample(p, vf, vo);
    }
    G.sAux[j] = make_float4(vf.x, vf.y, vf.z, d_wetAt(P, G, p));
  }
}
// sprayStep (129-181): every point moves by its kind; gone when its time is up,
// when it hits an object or leaves the box
__global__ void k_spray_step(const DParams* __restrict__ PP, const DCols* __restrict__ C, DGrid G) {
  const DParams& P = *PP;
  const uint32_t used = d_wwUsed(P, G);
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < used; j += gridDim.x * blockDim.x) {
    float4 sp = G.sPos[j];
    if (sp.w <= 0.0f) continue;
    float4 sv = G.sVel[j];
    float3 v = c3(sv);
    float kw = sv.w;
    uint32_t kind = st_f2u_sat(kw + 0.5f);
    float tk = (kw - float(kind)) / 0.4f * WWHOLD;
    const bool settled = tk >= WWHOLD * 0.999f;
    const uint32_t k0 = kind;
    float4 aux = G.sAux[j];
    float wet = aux.w;
    float3 vf = c3(aux);
    float life = sp.w;
    uint32_t h0 = j * 2654435761u + P.seed * 974711u;
    if (kind == 0u && v.y < vf.y && wet > 0.55f && tk > 0.05f) {
      kind = 2u;
      life = P.wwFoamLife * (0.6f + 0.8f * d_hash1(h0));
      v = vf;
    } else if (kind == 1u && wet > 0.55f) {
      life = 0.0f;
    } else if (kind == 2u && wet > 0.97f && settled) {
      kind = 3u;
    } else if (kind == 2u && wet < 0.04f && settled) {
      kind = 0u;
      life = fminf(life, P.sprLife);
    } else if (kind == 3u && wet < 0.6f && settled) {
      kind = 2u;
      life = fmaxf(life, 0.5f * P.wwFoamLife);
    }
    if (kind != k0) tk = 0.0f;
    else tk = fminf(tk + P.dt, WWHOLD);
    float3 p = c3(sp);
    if (kind <= 1u) {
      const bool mist = kind == 1u;
      float3 gust = make_float3(0.0f, 0.0f, 0.0f);
      if (P.q7 > 0.0f) {
        float3 gp = f3mul(f3mul(p, 0.55f), P.vs);
        float tt = P.q8 * 0.6f;
        gust = f3sub(f3mul(make_float3(d_vnz(f3add(gp, make_float3(tt, 0.0f, 0.0f))), d_vnz(f3add(gp, make_float3(17.3f, tt, 4.1f))),
                                       d_vnz(f3add(gp, make_float3(5.7f, 9.2f, tt)))),
                           2.0f),
                     make_float3(1.0f, 1.0f, 1.0f));
      }
      float3 base = make_float3(P.q1, P.q2, P.q3);
      if (P.airOn != 0u) base = d_airAt(P, G.air, p);
      float3 airv = f3add(base, f3mul(f3mul(gust, P.q7), mist ? 2.5f : 1.0f));
      v.y = v.y - P.gravity * P.dt;
      float t = 1.0f - st_expf(-(mist ? P.mistDrag : P.sprDrag) * P.dt);
      v = f3add(v, f3mul(f3sub(airv, v), t));
      life = life - P.dt;
    } else if (kind == 2u) {
      float vy = vf.y + st_clampf((wet - 0.5f) * 4.0f, -1.0f, 1.0f) * P.h * 4.0f;
      float t = 1.0f - st_expf(-10.0f * P.dt);
      v = make_float3(st_mixf(v.x, vf.x, t), st_mixf(v.y, vy, t), st_mixf(v.z, vf.z, t));
      life = life - P.dt;
    } else {
      float3 tv = f3add(vf, make_float3(0.0f, P.wwRise, 0.0f));
      float t = 1.0f - st_expf(-6.0f * P.dt);
      v = make_float3(st_mixf(v.x, tv.x, t), st_mixf(v.y, tv.y, t), st_mixf(v.z, tv.z, t));
      life = life - 0.3f * P.dt;
    }
    float3 np = f3add(p, f3mul(v, P.dt));
    if (np.x < P.x0 || np.x > P.x1 || np.y < P.y0 || np.y > P.y1 || np.z < P.z0 || np.z > P.z1) life = 0.0f;
    else if (P.ncols > 0u && d_colAt(P, C, G.sdf, np) >= 0) life = 0.0f;
    else if (P.ocean != 0u && kind >= 2u && d_layerW(P, np) > 0.6f) life = 0.0f;  // foam and bubbles end in the ocean's boundary
    G.sPos[j] = make_float4(np.x, np.y, np.z, life);
    G.sVel[j] = make_float4(v.x, v.y, v.z, float(kind) + 0.4f * tk / WWHOLD);
  }
}
// sprayEmit (98-127), the emission test and count of each particle: the count
// goes to sprCnt at the particle's id (scanned next), and the per-particle
// values the emission needs are recomputed by k_spray_emit
__device__ __forceinline__ uint32_t d_sprayCount(const DParams& P, const DGrid& G, float4 q, float4 vw, uint32_t id, float& rho,
                                                 float& fast, float& spd) {
  if (P.ocean != 0u && d_layerW(P, c3(q)) > 0.0f) return 0u;  // none from the water easing into the waves
  float3 v = c3(vw);
  float foam = vw.w;
  spd = f3len(v);
  fast = st_clampf((spd * P.vs - P.sprMin) / (2.0f * P.sprMin), 0.0f, 1.0f);
  int32_t bi = st_f2i_sat(floorf((q.x - P.x0) * P.ih - 0.5f)), bj = st_f2i_sat(floorf((q.y - P.y0) * P.ih - 0.5f)),
          bk = st_f2i_sat(floorf((q.z - P.z0) * P.ih - 0.5f));
  rho = 0.0f;
  for (int32_t dz = 0; dz <= 1; dz++)
    for (int32_t dy = 0; dy <= 1; dy++)
      for (int32_t dx = 0; dx <= 1; dx++) rho = rho + d_wetCell(P, G, bi + dx, bj + dy, bk + dz);
  rho = rho / 8.0f;
  float crest = (1.0f - st_smoothstepf(0.3f, 0.75f, rho)) * fast * (0.3f + 0.7f * foam);
  float churn = (1.0f - st_smoothstepf(0.8f, 1.0f, rho)) * st_clampf((foam - 0.55f) / 0.45f, 0.0f, 1.0f) *
                st_smoothstepf(0.5f, 1.5f, spd * P.vs);
  float e = (P.wwCrest * crest + P.wwImpact * churn) * st_smoothstepf(0.1f, 0.5f, P.q8);
  if (e <= 0.0f) return 0u;
  uint32_t h0 = id * 747796405u + P.seed * 2891336453u;
  return min(st_f2u_sat(floorf(P.sprRate * e * P.dt + d_hash1(h0))), 4u);
}
__global__ void k_spray_count(const DParams* __restrict__ PP, DGrid G) {
  const DParams& P = *PP;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < P.n; i += gridDim.x * blockDim.x) {
    float4 q = G.pos[i];
    uint32_t id = G.pid[i], k = 0u;
    if (q.y >= ST_DEADY) {
      float rho, fast, spd;
      k = d_sprayCount(P, G, q, G.vel[i], id, rho, fast, spd);
    }
    G.sprCnt[id] = k;
  }
}
__global__ void k_spray_emit(const DParams* __restrict__ PP, DGrid G) {
  const DParams& P = *PP;
  const uint32_t first = G.sctl[0];  // k_spray_advance moves it on after this kernel
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < P.n; i += gridDim.x * blockDim.x) {
    float4 q = G.pos[i];
    if (q.y < ST_DEADY) continue;
    uint32_t id = G.pid[i];
    float rho, fast, spd;
    float4 vw = G.vel[i];
    uint32_t k = d_sprayCount(P, G, q, vw, id, rho, fast, spd);
    if (!k) continue;
    float3 v = c3(vw);
    uint32_t h0 = id * 747796405u + P.seed * 2891336453u, e0 = G.sprOff[id];
    for (uint32_t m = 0; m < k; m++) {
      // the e-th point thrown off this substep takes slot (first + e) mod smax;
      // past smax points in one substep the ring has come round to slots this
      // substep already filled, which hold living points
      if (e0 + m >= P.smax) break;
      uint32_t hh = h0 ^ ((m + 1u) * 0x9E3779B9u);
      float z = d_hash1(hh + 1u) * 2.0f - 1.0f, a = d_hash1(hh + 2u) * 6.2831853f, s = sqrtf(fmaxf(0.0f, 1.0f - z * z));
      float sa, ca;
      st_sincosf(a, &sa, &ca);
      float3 dir = make_float3(s * ca, s * sa, z);
      bool under = rho > 0.7f && d_hash1(hh + 7u) < 0.7f;
      bool mist = !under && d_hash1(hh + 3u) < P.mistFrac * fast;
      float3 nv = f3add(f3mul(v, 0.9f + 0.2f * d_hash1(hh + 4u)), f3mul(f3mul(dir, spd), mist ? 0.2f : 0.06f));
      float life = P.sprLife * (0.6f + 0.8f * d_hash1(hh + 5u)) * (mist ? 1.6f : 1.0f);
      float kind = mist ? 1.0f : 0.0f;
      if (under) {
        nv = f3add(v, f3mul(f3mul(dir, spd), 0.05f));
        life = P.wwFoamLife * (0.6f + 0.8f * d_hash1(hh + 5u));
        kind = 3.0f;
      }
      uint32_t slot = (first + e0 + m) % P.smax;
      if (G.sPos[slot].w > 0.0f) continue;
      float3 np = f3add(f3add(c3(q), f3mul(nv, P.dt)), f3mul(f3mul(f3mul(dir, P.h), 0.3f), d_hash1(hh + 6u)));
      G.sPos[slot] = make_float4(np.x, np.y, np.z, life);
      G.sVel[slot] = make_float4(nv.x, nv.y, nv.z, kind);
    }
  }
}
// the ring counter moves on by every point asked for, placed or not
__global__ void k_spray_advance(const DParams* __restrict__ PP, DGrid G) {
  if (threadIdx.x == 0 && blockIdx.x == 0) G.sctl[0] = G.sctl[0] + G.sprOff[PP->n];
}

// dripStep (wgsl-whitewater.js:256-282): wet objects drip.  A sample over a
// collider's surface is wet while water is at it and dries over dripDry seconds
// once out; while drying it lets go of drops (more from surfaces facing down),
// which fly as spray with the object's velocity.  The wetness is updated and
// each sample's drop decided here; the drops take consecutive ring slots in
// sample order (k_drip_emit).
__device__ __forceinline__ void d_dripGeom(const DCols* __restrict__ C, float4 a, float4 b, uint32_t k, const DParams& P, float3& wp,
                                           float3& wn, float3& probe) {
  float3 r0 = c3(C->c[k * 8u + 2u]), r1 = c3(C->c[k * 8u + 3u]), r2 = c3(C->c[k * 8u + 4u]);
  wp = f3add(f3add(f3add(c3(C->c[k * 8u + 5u]), f3mul(r0, a.x)), f3mul(r1, a.y)), f3mul(r2, a.z));
  float3 n = f3add(f3add(f3mul(r0, b.x), f3mul(r1, b.y)), f3mul(r2, b.z));
  wn = f3div(n, f3len(n));
  probe = f3add(wp, f3mul(f3mul(wn, P.h), 0.75f));
}
__global__ void k_drip_flags(const DParams* __restrict__ PP, const DCols* __restrict__ C, DGrid G) {
  const DParams& P = *PP;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < P.ndrip; s += gridDim.x * blockDim.x) {
    G.dripFlag[s] = 0u;
    float4 a = G.drip[s * 2u], b = G.drip[s * 2u + 1u];
    uint32_t k = st_f2u_sat(a.w);
    if (k >= P.ncols) continue;
    float3 wp, wn, probe;
    d_dripGeom(C, a, b, k, P, wp, wn, probe);
    float here = d_wetAt(P, G, probe);
    if (here > 0.4f) {
      G.drip[s * 2u + 1u].w = 1.0f;
      continue;
    }
    float wetness = fmaxf(0.0f, b.w - P.dt / fmaxf(P.dripDry, 0.05f));
    G.drip[s * 2u + 1u].w = wetness;
    if (wetness <= 0.0f || here > 0.1f || P.smax == 0u || P.dripRate <= 0.0f) continue;
    float down = st_clampf(-wn.y, 0.0f, 1.0f);
    float rate = P.dripRate * wetness * wetness * (0.15f + 0.85f * down);
    uint32_t h0 = s * 747796405u + P.seed * 2891336453u + 17u;
    if (d_hash1(h0) >= rate * P.dt) continue;
    G.dripFlag[s] = 1u;
  }
}
__global__ void k_drip_emit(const DParams* __restrict__ PP, const DCols* __restrict__ C, DGrid G) {
  const DParams& P = *PP;
  const uint32_t first = G.sctl[0];
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < P.ndrip; s += gridDim.x * blockDim.x) {
    if (!G.dripFlag[s]) continue;
    const uint32_t e = G.dripOff[s];
    if (e >= P.smax) continue;  // the ring came round to drops placed this substep
    float4 a = G.drip[s * 2u], b = G.drip[s * 2u + 1u];
    uint32_t k = st_f2u_sat(a.w);
    float3 wp, wn, probe;
    d_dripGeom(C, a, b, k, P, wp, wn, probe);
    uint32_t slot = (first + e) % P.smax;
    if (G.sPos[slot].w > 0.0f) continue;
    uint32_t h0 = s * 747796405u + P.seed * 2891336453u + 17u;
    float3 v = f3add(d_colSurfVel(C, k, wp), f3mul(f3mul(f3mul(wn, P.h), 2.0f), d_hash1(h0 + 1u)));
    G.sPos[slot] = make_float4(probe.x, probe.y, probe.z, P.sprLife * 2.0f);
    G.sVel[slot] = make_float4(v.x, v.y, v.z, 0.0f);
  }
}
__global__ void k_drip_advance(const DParams* __restrict__ PP, DGrid G) {
  if (threadIdx.x == 0 && blockIdx.x == 0) G.sctl[0] = G.sctl[0] + G.dripOff[PP->ndrip];
}

void launch_spray(hipStream_t st, const LaunchCtx& L, uint32_t airN, bool air, bool emit, uint32_t ndrip) {
  const uint32_t pg = std::max<uint32_t>(1u, std::min<uint32_t>(gridFor(L.n, WG), 8192u));
  const uint32_t sg = std::max<uint32_t>(1u, std::min<uint32_t>(gridFor(L.smax, WG), 8192u));
  if (air && airN) {
    launch_fill_u32(st, (uint32_t*)L.G.airAcc, size_t(airN) * 4u, 0u);
    const uint32_t ag = std::max<uint32_t>(1u, std::min<uint32_t>(gridFor(airN, WG), 4096u));
    if (L.n) hipLaunchKernelGGL(k_air_splat_w, dim3(pg), dim3(WG), 0, st, L.P, L.G);
    hipLaunchKernelGGL(k_air_splat_s, dim3(sg), dim3(WG), 0, st, L.P, L.G);
    hipLaunchKernelGGL(k_air_update, dim3(ag), dim3(WG), 0, st, L.P, L.C, L.G);
    hipLaunchKernelGGL(k_air_copy, dim3(ag), dim3(WG), 0, st, L.P, L.G);
  }
  hipLaunchKernelGGL(k_ww_sample, dim3(sg), dim3(WG), 0, st, L.P, L.G);
  hipLaunchKernelGGL(k_spray_step, dim3(sg), dim3(WG), 0, st, L.P, L.C, L.G);
  if (emit && L.n) {
    hipLaunchKernelGGL(k_spray_count, dim3(pg), dim3(WG), 0, st, L.P, L.G);
    launch_scan(st, L.G.sprCnt, L.G.sprOff, L.n, L.G.sprSums, nullptr);
    hipLaunchKernelGGL(k_spray_emit, dim3(pg), dim3(WG), 0, st, L.P, L.G);
    hipLaunchKernelGGL(k_spray_advance, dim3(1), dim3(64), 0, st, L.P, L.G);
  }
  if (ndrip) {
    const uint32_t dg = std::max<uint32_t>(1u, std::min<uint32_t>(gridFor(ndrip, WG), 1024u));
    hipLaunchKernelGGL(k_drip_flags, dim3(dg), dim3(WG), 0, st, L.P, L.C, L.G);
    launch_scan(st, L.G.dripFlag, L.G.dripOff, ndrip, L.G.dripSums, nullptr);
    hipLaunchKernelGGL(k_drip_emit, dim3(dg), dim3(WG), 0, st, L.P, L.C, L.G);
    hipLaunchKernelGGL(k_drip_advance, dim3(1), dim3(64), 0, st, L.P, L.G);
  }
}

// ---------------------------------------------------------------------------
// The live view (sim.js:364-404, 504-537; wgsl-whitewater.js:183-253): a sample
// of the water and of the whitewater packed into 32-bit words in the cache
// frame format -- a 4-word header, then per group of four water points six words
// of 16-bit positions, a word of speed bytes and a word of white-water bytes,
// then per group of four whitewater points six words of positions and four
// words of 32-bit codes.  The water sample is every stride-th particle in id
// order (the reference's buffer order); the whitewater sample is every s-th
// slot of those handed out, the living ones packed together in slot order
// (where the reference's atomics pack them in arrival order).

// packSpray (wgsl-whitewater.js:184-195)
__device__ __forceinline__ uint32_t d_packSpray(float3 v, uint32_t kind, float lifeFrac) {
  float s = f3len(v);
  uint32_t code = 0x80000000u | (min(st_f2u_sat(s * 4.0f), 255u) << 21);
  if (kind == 1u) code |= 0x40000000u;
  else if (kind == 2u) code |= 0x20000000u;
  else if (kind == 3u) code |= 0x60000000u;
  if (kind >= 2u) return code | st_f2u_sat(st_clampf(lifeFrac, 0.0f, 1.0f) * 1023.0f);
  float3 n = make_float3(0.0f, 1.0f, 0.0f);
  if (s > 1e-6f) n = f3div(v, s);
  float l1 = fabsf(n.x) + fabsf(n.y) + fabsf(n.z);
  float px = n.x / l1, py = n.y / l1;
  if (n.z < 0.0f) {
    float qx = (1.0f - fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f);
    float qy = (1.0f - fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f);
    px = qx;
    py = qy;
  }
  uint32_t u = st_f2u_sat(st_clampf(rintf((px * 0.5f + 0.5f) * 1023.0f), 0.0f, 1023.0f));
  uint32_t w = st_f2u_sat(st_clampf(rintf((py * 0.5f + 0.5f) * 1023.0f), 0.0f, 1023.0f));
  return code | (w << 10) | u;
}
// sprayDisp (wgsl-whitewater.js:201-211): which sample slots hold a living point
__global__ void k_view_spray_flags(const DParams* __restrict__ PP, DGrid G, uint32_t sdMax, uint32_t* __restrict__ flag) {
  const DParams& P = *PP;
  const uint32_t used = d_wwUsed(P, G), stride = max(1u, (used + sdMax - 1u) / sdMax);
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < sdMax; k += gridDim.x * blockDim.x) {
    uint32_t j = k * stride;
    flag[k] = (j < used && G.sPos[j].w > 0.0f) ? 1u : 0u;
  }
}
__global__ void k_view_spray_put(const DParams* __restrict__ PP, DGrid G, uint32_t sdMax, const uint32_t* __restrict__ flag,
                                 const uint32_t* __restrict__ off, float4* __restrict__ disp, float foamLife) {
  const DParams& P = *PP;
  const uint32_t used = d_wwUsed(P, G), stride = max(1u, (used + sdMax - 1u) / sdMax);
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < sdMax; k += gridDim.x * blockDim.x) {
    if (!flag[k]) continue;
    uint32_t j = k * stride;
    float4 sp = G.sPos[j], sv = G.sVel[j];
    uint32_t code = d_packSpray(c3(sv), st_f2u_sat(sv.w + 0.5f), sp.w / fmaxf(foamLife, 0.01f));
    disp[off[k]] = make_float4(sp.x, sp.y, sp.z, __uint_as_float(code));
  }
}
__device__ __forceinline__ uint3 d_q16(const ViewParams& V, float3 p) {
  return make_uint3(st_f2u_sat(st_clampf(rintf((p.x - V.x0) * V.kx), 0.0f, 65535.0f)),
                    st_f2u_sat(st_clampf(rintf((p.y - V.y0) * V.ky), 0.0f, 65535.0f)),
                    st_f2u_sat(st_clampf(rintf((p.z - V.z0) * V.kz), 0.0f, 65535.0f)));
}
__device__ __forceinline__ void d_put6(uint32_t* out, uint32_t o, const uint3* q) {
  out[o] = q[0].x | (q[0].y << 16);
  out[o + 1u] = q[0].z | (q[1].x << 16);
  out[o + 2u] = q[1].y | (q[1].z << 16);
  out[o + 3u] = q[2].x | (q[2].y << 16);
  out[o + 4u] = q[2].z | (q[3].x << 16);
  out[o + 5u] = q[3].y | (q[3].z << 16);
}
// packDisp (wgsl-whitewater.js:219-253) for the water and whitewater parts;
// water points come from the id-ordered copy `wat` (every srcStride-th one)
__global__ void k_pack_view(const ViewParams V, const float4* __restrict__ wat, const float4* __restrict__ sdisp,
                            const uint32_t* __restrict__ nSp, uint32_t* __restrict__ out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t gW = (V.nW + 3u) / 4u, gS = (V.nS + 3u) / 4u;
  const uint32_t nS = min(*nSp, V.nS);
  if (t == 0u) {
    out[0] = nS;
    out[1] = 0u;
    out[2] = 0u;
    out[3] = 0u;
  }
  uint3 q[4];
  if (t < gW) {
    uint32_t b0 = 0u, b1 = 0u;
    for (uint32_t m = 0; m < 4u; m++) {
      const uint32_t i = t * 4u + m;
      float4 v = make_float4(0.0f, -1.0e6f, 0.0f, 0.0f);
      if (i < V.nW) v = wat[size_t(i) * V.srcStride];
      uint32_t s = 255u, f = 0u;
      q[m] = make_uint3(0u, 0u, 0u);
      if (v.y > ST_DEADY) {
        q[m] = d_q16(V, c3(v));
        float lv = floorf(v.w / 64.0f);
        s = st_f2u_sat(st_clampf((v.w - lv * 64.0f) * 16.0f * V.vs, 0.0f, 191.0f));
        f = st_f2u_sat(st_clampf(lv, 0.0f, 15.0f)) * 17u;
      }
      b0 |= s << (8u * m);
      b1 |= f << (8u * m);
    }
    d_put6(out, V.oWp + t * 6u, q);
    out[V.oWs + t] = b0;
    out[V.oWf + t] = b1;
    return;
  }
  const uint32_t ts = t - gW;
  if (ts < gS) {
    for (uint32_t m = 0; m < 4u; m++) {
      const uint32_t k = ts * 4u + m;
      uint32_t code = 0u;
      q[m] = make_uint3(0u, 0u, 0u);
      if (k < nS) {
        float4 v = sdisp[k];
        if (v.y > ST_DEADY) {
          q[m] = d_q16(V, c3(v));
          code = __float_as_uint(v.w);
        }
      }
      out[V.oSk + k] = code;
    }
    d_put6(out, V.oSp + ts * 6u, q);
  }
}

// deep water points (deepClear/deepPick/deepFin, wgsl-whitewater.js:279-297), narrow
// band only: the centre of every fluid cell in the domain that holds no particle,
// w = 4096 + min(|kept grid velocity there|, 63).  Above DEEP_MAX such cells
// (counted by the previous read, `prev`) a cell is kept when hash1(c 2654435761 +
// 7) prev < DEEP_MAX.  The points are written in active-cell order (slot list,
// then cell) by an exclusive scan of the kept flags, where the reference appends
// them with an atomic counter in arrival order; at most DEEP_MAX are written.
__global__ void __launch_bounds__(WG) k_deep_flags(const DParams* __restrict__ PP, DGrid G, uint32_t prev, uint32_t* __restrict__ flag,
                                                   uint32_t* __restrict__ all) {
  const DParams& P = *PP;
  const uint32_t tot = G.ctl[4] * 512u;  // a multiple of the block size: whole waves below
  for (uint32_t t = blockIdx.x * WG + threadIdx.x; t < tot; t += gridDim.x * WG) {
    const uint32_t s = G.slotList[t >> 9], l = t & 511u, c = s * 512u + l;
    const int3 at = d_slotCoord(P, G.slotTile, s, l);
    const bool deep = d_inDom(P, at) && (G.ctype[c] & 0xffu) == 1u && G.cnt[c] <= 0;
    bool keep = deep;
    if (deep && prev > DEEP_MAX && d_hash1(c * 2654435761u + 7u) * float(prev) >= float(DEEP_MAX)) keep = false;
    flag[t] = keep ? 1u : 0u;
    const uint64_t m = __ballot(deep);
    if ((threadIdx.x & 63u) == 0u && m) atomicAdd(all, uint32_t(__popcll(m)));
  }
}
__global__ void __launch_bounds__(WG) k_deep_emit(const DParams* __restrict__ PP, DGrid G, const uint32_t* __restrict__ flag,
                                                  const uint32_t* __restrict__ off, float4* __restrict__ out) {
  const DParams& P = *PP;
  const uint32_t tot = G.ctl[4] * 512u;
  for (uint32_t t = blockIdx.x * WG + threadIdx.x; t < tot; t += gridDim.x * WG) {
    if (!flag[t] || off[t] >= DEEP_MAX) continue;
    const uint32_t s = G.slotList[t >> 9], l = t & 511u;
    const int3 at = d_slotCoord(P, G.slotTile, s, l);
    const float3 x = d_cellCenter(P, at.x, at.y, at.z);
    out[off[t]] = make_float4(x.x, x.y, x.z, 4096.0f + fminf(f3len(d_velGVec(P, G, x)), 63.0f));
  }
}
void launch_deep(hipStream_t st, const LaunchCtx& L, uint32_t ncell, uint32_t prev, uint32_t* flag, uint32_t* off, uint32_t* sums,
                 uint32_t* all, float4* out) {
  const uint32_t g = std::max<uint32_t>(1u, std::min<uint32_t>(gridFor(ncell, WG), 8192u));
  launch_fill_u32(st, all, 1, 0u);
  launch_fill_u32(st, flag, ncell, 0u);  // cells past the slots in use stay unflagged
  hipLaunchKernelGGL(k_deep_flags, dim3(g), dim3(WG), 0, st, L.P, L.G, prev, flag, all);
  launch_scan(st, flag, off, ncell, sums, nullptr);
  hipLaunchKernelGGL(k_deep_emit, dim3(g), dim3(WG), 0, st, L.P, L.G, flag, off, out);
}

void launch_pack_view(hipStream_t st, const LaunchCtx& L, const ViewParams& V, uint32_t sdMax, uint32_t* flag, uint32_t* off,
                      uint32_t* sums, float4* sdisp, uint32_t* out) {
  if (V.n) launch_unsort(st, L.G.pos, L.G.pid, L.G.scratch4, V.n);
  if (sdMax) {
    const uint32_t g = std::max<uint32_t>(1u, std::min<uint32_t>(gridFor(sdMax, WG), 4096u));
    hipLaunchKernelGGL(k_view_spray_flags, dim3(g), dim3(WG), 0, st, L.P, L.G, sdMax, flag);
    launch_scan(st, flag, off, sdMax, sums, nullptr);
    hipLaunchKernelGGL(k_view_spray_put, dim3(g), dim3(WG), 0, st, L.P, L.G, sdMax, flag, off, sdisp, V.foamLife);
  } else {
    launch_fill_u32(st, off, 1, 0u);
  }
  const uint32_t groups = (V.nW + 3u) / 4u + (V.nS + 3u) / 4u;
  hipLaunchKernelGGL(k_pack_view, dim3(std::max<uint32_t>(1u, gridFor(groups, WG))), dim3(WG), 0, st, V, L.G.scratch4, sdisp,
                     sdMax ? off + sdMax : off, out);
}

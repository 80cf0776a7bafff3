// flip_adjoint.hip - hand-written adjoint kernels of the FLIP substep.
//
// The backward pass of one substep at fixed topology: the tile set, cell types,
// particle counts and every branch (clamps, wall contacts) are those the forward
// substep took, and the fixed-point sums of p2g are differentiated as the real
// sums they round (the rounding is a step function; its derivative is taken as
// one).  Stages, reversed:
//   g2p       per particle: gradient of (position, velocity) after the substep ->
//             velocity before it (FLIP term), position (advection, wall clamp,
//             trilinear weights) and the new/old face velocities it sampled
//             (scattered with f32 atomics; order does not matter at the f32 bar)
//   project   face gradient passes through; the pressure gets -+dt/h of it
//   SOR       the red-black sweeps transposed, in reverse order (black then red
//             per iteration): relaxing colour c, p_c <- (1-w) p_c + w/n (rhs_c +
//             sum of fluid neighbours), has the transpose g_c <- (1-w) g_c,
//             g_rhs_c += w/n g_c, g_m += w/n_c g_c for every fluid neighbour m
//   (the gradient left on the warm start goes to the previous substep's pressure)
//   divergence rhs = (tgt - div) h^2/dt  ->  face gradients
//   normFaces v = V / W on faces with no solid side (old velocity the same,
//             new minus g dt on y faces)  ->  gradients of the sums V and W
//   p2g       per particle: V = sum w v, W = sum w over its 8 faces per axis
//             -> velocity and position gradients
// Colliders (at most ADJ_MAX_COLS) take part through the particle push of g2p,
// p += n (q0 - phi) with phi and n the trilinear distance and its unit gradient
// (so their derivatives need the trilinear Hessian), and through the solid faces
// of normFaces, whose velocity is v + w x (x - c); both give the collider's
// centre, linear and angular velocity a gradient (dL/dc, dL/dv, dL/dw).
// References: p2g wgsl-solver.js:16-40, normFaces 75-105, divergence 106-123,
// relax 124-140, project 165-180, g2p 208-275, collideP 185-197.
#include "flip_adjoint.h"
#include "grid_barrier.h"

#include <algorithm>

#define AWG 256

// sparse view of a recorded substep for the addressing helpers
struct AView {
  const DParams& P;
  const AdjTape& T;
  __device__ __forceinline__ uint32_t face(uint32_t axis, int32_t i, int32_t j, int32_t k) const {
    return d_faceIndex(P, T.tmap, axis, i, j, k);
  }
  __device__ __forceinline__ uint32_t type(int32_t i, int32_t j, int32_t k) const { return d_ctypeOf(P, T.tmap, T.ctype, i, j, k); }
  __device__ __forceinline__ uint32_t cell(int32_t i, int32_t j, int32_t k) const { return d_cellIdx(P, T.tmap, i, j, k); }
};

__device__ __forceinline__ float3 axisGrid(const DParams& P, float3 p, uint32_t axis) {
  float3 g = make_float3((p.x - P.x0) * P.ih, (p.y - P.y0) * P.ih, (p.z - P.z0) * P.ih);
  if (axis == 0) { g.y -= 0.5f; g.z -= 0.5f; }
  else if (axis == 1) { g.x -= 0.5f; g.z -= 0.5f; }
  else { g.x -= 0.5f; g.y -= 0.5f; }
  return g;
}

// ---------------------------------------------------------------------------
// collider push (collideP, wgsl-solver.js:185-197; sdfEval, gpu.js:6-30) reversed.
// In the collider's grid coordinates g = (R^T (p - c) - origin) / cell the
// distance is trilinear inside the grid, with gradient graw (the forward's
// unnormalised normal) and a Hessian with a zero diagonal; outside it, phi =
// trilinear(clamped g) + |g - clamped g| cell and the normal is the unit offset.
// For the push p' = p + wn (q0 - phi) and a gradient u of p', returns the
// gradient of p; that of the centre is minus the part beyond u.
__device__ __forceinline__ float3 adj_push(const DParams& P, const DCols* __restrict__ C, const float* __restrict__ sdf, uint32_t k,
                                           float3 p, float3 u) {
  const float4 a = C->c[k * 8u], b = C->c[k * 8u + 1u];
  const float3 r0 = c3(C->c[k * 8u + 2u]), r1 = c3(C->c[k * 8u + 3u]), r2 = c3(C->c[k * 8u + 4u]);
  const float3 d = f3sub(p, c3(C->c[k * 8u + 5u]));
  const float3 lp = make_float3(f3dot(r0, d), f3dot(r1, d), f3dot(r2, d));
  const float3 g = f3div(f3sub(lp, make_float3(b.x, b.y, b.z)), b.w);
  const float e = 1.0001f;
  const float3 gc = make_float3(st_clampf(g.x, 0.0f, a.y - e), st_clampf(g.y, 0.0f, a.z - e), st_clampf(g.z, 0.0f, a.w - e));
  const float3 fl = make_float3(floorf(gc.x), floorf(gc.y), floorf(gc.z));
  const float3 f = f3sub(gc, fl);
  const uint32_t snx = st_f2u_sat(a.y), sny = st_f2u_sat(a.z);
  const uint32_t base = st_f2u_sat(a.x) + st_f2u_sat(fl.x) + snx * (st_f2u_sat(fl.y) + sny * st_f2u_sat(fl.z));
  const uint32_t s = snx * sny;
  const float v000 = sdf[base], v100 = sdf[base + 1u], v010 = sdf[base + snx], v110 = sdf[base + snx + 1u];
  const float v001 = sdf[base + s], v101 = sdf[base + s + 1u], v011 = sdf[base + s + snx], v111 = sdf[base + s + snx + 1u];
  const float c00 = st_mixf(v000, v100, f.x), c10 = st_mixf(v010, v110, f.x), c01 = st_mixf(v001, v101, f.x), c11 = st_mixf(v011, v111, f.x);
  const float c0 = st_mixf(c00, c10, f.y), c1 = st_mixf(c01, c11, f.y);
  float phi = st_mixf(c0, c1, f.z);
  const float3 graw = make_float3(st_mixf(st_mixf(v100 - v000, v110 - v010, f.y), st_mixf(v101 - v001, v111 - v011, f.y), f.z),
                                  st_mixf(c10 - c00, c11 - c01, f.z), c1 - c0);
  const float3 out = f3sub(g, gc);
  float3 nl, dphi;   // unit normal and dphi/dg, grid coordinates
  float J[3][3];     // d nl / d g
  if (f3dot(out, out) > 0.0f) {
    const float l = f3len(out);
    phi = phi + l * b.w;
    nl = f3div(out, l);
    const float m[3] = {g.x != gc.x ? 1.f : 0.f, g.y != gc.y ? 1.f : 0.f, g.z != gc.z ? 1.f : 0.f};
    const float n3[3] = {nl.x, nl.y, nl.z}, gr[3] = {graw.x, graw.y, graw.z};
    float dp[3];
    for (int i = 0; i < 3; i++) {
      dp[i] = m[i] > 0.f ? n3[i] * b.w : gr[i];
      for (int j = 0; j < 3; j++) J[i][j] = ((i == j ? 1.f : 0.f) - n3[i] * n3[j]) / l * m[j];
    }
    dphi = make_float3(dp[0], dp[1], dp[2]);
  } else {
    const float hxy = st_mixf((v110 - v010) - (v100 - v000), (v111 - v011) - (v101 - v001), f.z);
    const float hxz = st_mixf(v101 - v001, v111 - v011, f.y) - st_mixf(v100 - v000, v110 - v010, f.y);
    const float hyz = (c11 - c01) - (c10 - c00);
    const float H[3][3] = {{0.f, hxy, hxz}, {hxy, 0.f, hyz}, {hxz, hyz, 0.f}};
    const float l = f3len(graw), ld = fmaxf(l, 1e-9f);
    nl = f3div(graw, ld);
    const float n3[3] = {nl.x, nl.y, nl.z};
    const bool proj = l > 1e-9f;  // below it the forward divides by a constant
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        float v = 0.f;
        for (int q = 0; q < 3; q++) v += ((i == q ? 1.f : 0.f) - (proj ? n3[i] * n3[q] : 0.f)) * H[q][j];
        J[i][j] = v / ld;
      }
    dphi = graw;
  }
  const float3 wn = f3add(f3add(f3mul(r0, nl.x), f3mul(r1, nl.y)), f3mul(r2, nl.z));
  const float sh = P.q0 - phi;
  // (d wn / d p)^T u = R J^T R^T u / cell, d phi / d p = R dphi / cell (R rows r0..r2 map world to grid)
  const float3 Mu = make_float3(f3dot(r0, u), f3dot(r1, u), f3dot(r2, u));
  const float3 JtMu = make_float3(J[0][0] * Mu.x + J[1][0] * Mu.y + J[2][0] * Mu.z, J[0][1] * Mu.x + J[1][1] * Mu.y + J[2][1] * Mu.z,
                                  J[0][2] * Mu.x + J[1][2] * Mu.y + J[2][2] * Mu.z);
  const float3 ga = f3div(f3add(f3add(f3mul(r0, JtMu.x), f3mul(r1, JtMu.y)), f3mul(r2, JtMu.z)), b.w);
  const float3 dphiP = f3div(f3add(f3add(f3mul(r0, dphi.x), f3mul(r1, dphi.y)), f3mul(r2, dphi.z)), b.w);
  return f3add(u, f3sub(f3mul(ga, sh), f3mul(dphiP, f3dot(wn, u))));
}

// ---------------------------------------------------------------------------
// g2p (wgsl-solver.js:208-275) reversed.  Needs drag 0 and closed walls.
// One workgroup per tile bin (the bins the forward substep used): the face
// gradients of the tile's particles gather in a 10^3-cell LDS window (the tile
// plus one cell all round holds every face they sampled) and go to HBM once per
// tile, so particles of one tile do not queue on the same global addresses.

#define AWIN 10
#define AWIN3 1000

// per particle: the gradient of its state before the substep from that after
// it; the face gradients go to add(axis, face, d new, d old)
template <class Fetch, class Add>
__device__ __forceinline__ void adj_g2p_particle(const DParams& P, const AdjTape& T, const AdjGrad& G, uint32_t i, bool storage,
                                                 float* sCol, Fetch fetch, Add add) {
  const float4 p4 = T.pos[i], v4 = T.vel[i];
  const uint32_t id = T.pid[i];
  const float3 p = c3(p4), v0 = c3(v4);
  // forward again: sampled new / old velocity, FLIP blend
  float3 vn = make_float3(v0.x, v0.y - P.gravity * P.dt, v0.z), vo = v0;
  if (storage) {
    float s_n[3], s_o[3];
    for (uint32_t a = 0; a < 3; a++) {
      const float3 g = axisGrid(P, p, a);
      const float3 f0 = make_float3(floorf(g.x), floorf(g.y), floorf(g.z));
      const float3 fr = f3sub(g, f0);
      const int3 i0 = make_int3(int32_t(f0.x), int32_t(f0.y), int32_t(f0.z));
      float sn = 0.f, so = 0.f;
      for (uint32_t c = 0; c < 8; c++) {
        const int32_t dx = c & 1, dy = (c >> 1) & 1, dz = (c >> 2) & 1;
        const float w = (dx ? fr.x : 1.f - fr.x) * (dy ? fr.y : 1.f - fr.y) * (dz ? fr.z : 1.f - fr.z);
        const float2 f = fetch(a, i0.x + dx, i0.y + dy, i0.z + dz);
        sn += w * f.x;
        so += w * f.y;
      }
      s_n[a] = sn;
      s_o[a] = so;
    }
    vn = make_float3(s_n[0], s_n[1], s_n[2]);
    vo = make_float3(s_o[0], s_o[1], s_o[2]);
  }
  const float3 vm = f3add(vn, f3mul(f3sub(v0, vo), P.flip));
  const float sp = f3len(vm);
  const bool capped = sp > P.vmax;
  const float3 vc = capped ? f3mul(vm, P.vmax / sp) : vm;
  const float3 pt = f3add(p, f3mul(vc, P.dt));
  // collider pushes in order (the positions they start from are kept), then the box clamp
  float3 q = pt;
  float3 qk[ADJ_MAX_COLS];
  uint32_t hit = 0u;
  const uint32_t ncol = min(P.ncols, uint32_t(ADJ_MAX_COLS));  // the record refuses more
#pragma unroll
  for (uint32_t k = 0; k < ADJ_MAX_COLS; k++) {  // unrolled: qk stays in registers
    if (k >= ncol || !d_colNear(P, T.C, k, q)) continue;
    float phi;
    float3 wn;
    d_sdfEval(T.C, T.sdf, k, q, phi, wn);
    if (phi < P.q0) {
      qk[k] = q;
      hit |= 1u << k;
      q = f3add(q, f3mul(wn, P.q0 - phi));
    }
  }
  const float m = P.margin;
  const float3 lo = make_float3(P.x0 + m, P.y0 + m, P.z0 + m), hi = make_float3(P.x1 - m, P.y1 - m, P.z1 - m);
  const float3 pc = make_float3(st_clampf(q.x, lo.x, hi.x), st_clampf(q.y, lo.y, hi.y), st_clampf(q.z, lo.z, hi.z));
  const float3 keep = make_float3(q.x > lo.x && q.x < hi.x ? 1.f : 0.f, q.y > lo.y && q.y < hi.y ? 1.f : 0.f,
                                  q.z > lo.z && q.z < hi.z ? 1.f : 0.f);
  const float3 dsp = f3sub(pc, pt);
  bool wall = false;
  float3 nrm = make_float3(0.f, 0.f, 0.f);
  float dl = 0.f;
  if (f3dot(dsp, dsp) > 1.0e-12f) {
    dl = f3len(dsp);
    nrm = f3div(dsp, dl);
    wall = f3dot(vc, nrm) < 0.f;
  }
  // backward: the contact removes the normal velocity, v' = vc - n (n . vc), with
  // n = dsp / |dsp| a function of the displacement the pushes and the clamp made
  const float3 gpc = c3(G.gPos[id]), gvo = c3(G.gVel[id]);
  float3 gvc = gvo, gdsp = make_float3(0.f, 0.f, 0.f);
  if (wall) {
    gvc = f3sub(gvo, f3mul(nrm, f3dot(nrm, gvo)));
    const float3 gn = f3mul(f3add(f3mul(gvo, f3dot(nrm, vc)), f3mul(vc, f3dot(gvo, nrm))), -1.f);
    gdsp = f3div(f3sub(gn, f3mul(nrm, f3dot(nrm, gn))), dl);
  }
  const float3 gpt0 = f3add(gpc, gdsp);
  float3 gq = make_float3(keep.x * gpt0.x, keep.y * gpt0.y, keep.z * gpt0.z);
#pragma unroll
  for (int k = ADJ_MAX_COLS - 1; k >= 0; k--) {
    if (!((hit >> k) & 1u)) continue;
    const float3 gb = adj_push(P, T.C, T.sdf, uint32_t(k), qk[k], gq);
    const float3 gcen = f3sub(gq, gb);  // the centre enters as p - c
    atomicAdd(&sCol[k * 3 + 0], gcen.x);
    atomicAdd(&sCol[k * 3 + 1], gcen.y);
    atomicAdd(&sCol[k * 3 + 2], gcen.z);
    gq = gb;
  }
  float3 gp = f3sub(gq, gdsp);
  gvc = f3add(gvc, f3mul(gp, P.dt));
  float3 gvm = gvc;
  if (capped) {
    const float3 u = f3div(vm, sp);
    gvm = f3mul(f3sub(gvc, f3mul(u, f3dot(u, gvc))), P.vmax / sp);
  }
  float3 gv0 = gvm;  // without storage vn = v0 - g dt and vo = v0
  if (storage) {
    gv0 = f3mul(gvm, P.flip);
    const float gvmA[3] = {gvm.x, gvm.y, gvm.z};
    for (uint32_t a = 0; a < 3; a++) {
      const float3 g = axisGrid(P, p, a);
      const float3 f0 = make_float3(floorf(g.x), floorf(g.y), floorf(g.z));
      const float3 fr = f3sub(g, f0);
      const int3 i0 = make_int3(int32_t(f0.x), int32_t(f0.y), int32_t(f0.z));
      const float gn = gvmA[a], go = -P.flip * gvmA[a];
      for (uint32_t c = 0; c < 8; c++) {
        const int32_t dx = c & 1, dy = (c >> 1) & 1, dz = (c >> 2) & 1;
        const float wx = dx ? fr.x : 1.f - fr.x, wy = dy ? fr.y : 1.f - fr.y, wz = dz ? fr.z : 1.f - fr.z;
        const float sx = dx ? P.ih : -P.ih, sy = dy ? P.ih : -P.ih, sz = dz ? P.ih : -P.ih;
        const float2 f = fetch(a, i0.x + dx, i0.y + dy, i0.z + dz);
        const float w = wx * wy * wz;
        add(a, i0.x + dx, i0.y + dy, i0.z + dz, w * gn, w * go);
        const float s = gn * f.x + go * f.y;
        gp = f3add(gp, make_float3(sx * wy * wz * s, wx * sy * wz * s, wx * wy * sz * s));
      }
    }
  }
  G.gPos2[id] = make_float4(gp.x, gp.y, gp.z, 0.f);
  G.gVel2[id] = make_float4(gv0.x, gv0.y, gv0.z, 0.f);
}

// face (axis, i, j, k) as faceIndex clamps it, relative to the window origin
__device__ __forceinline__ int32_t adj_window_face(const DParams& P, uint32_t a, int32_t i, int32_t j, int32_t k, int3 org) {
  const int32_t X = int32_t(P.nx), Y = int32_t(P.ny), Z = int32_t(P.nz);
  i = d_iclamp(i, 0, a == 0 ? X : X - 1);
  j = d_iclamp(j, 0, a == 1 ? Y : Y - 1);
  k = d_iclamp(k, 0, a == 2 ? Z : Z - 1);
  return (i - org.x) + AWIN * ((j - org.y) + AWIN * (k - org.z));
}

#define AG2P_WG 512
__global__ void __launch_bounds__(AG2P_WG) k_adj_g2p(AdjTape T, AdjGrad G) {
  const DParams& P = *T.P;
  const AView V{P, T};
  __shared__ float2 sF[3][AWIN3];  // (new, old) face velocity of the window
  __shared__ float sN[3][AWIN3];   // gradient gathered for the new face velocity
  __shared__ float sO[3][AWIN3];   // and for the old one
  __shared__ float sCol[ADJ_MAX_COLS * 3];  // centre gradients of the colliders
  if (threadIdx.x < ADJ_MAX_COLS * 3) sCol[threadIdx.x] = 0.f;
  __syncthreads();
  const int32_t gx = int32_t(P.tx) * 8, gy = int32_t(P.ty) * 8, gz = int32_t(P.tz) * 8;
  auto none = [](uint32_t, int32_t, int32_t, int32_t) { return make_float2(0.f, 0.f); };
  auto drop = [](uint32_t, int32_t, int32_t, int32_t, float, float) {};
  for (uint32_t t = blockIdx.x; t <= P.ntiles; t += gridDim.x) {
    const uint32_t beg = T.binStart[t], end = T.binStart[t + 1];
    if (beg == end) continue;
    if (t == P.ntiles) {  // dead particles: no gradient
      for (uint32_t k = beg + threadIdx.x; k < end; k += AG2P_WG) {
        const uint32_t id = T.pid[T.binIdx[k]];
        G.gPos2[id] = make_float4(0.f, 0.f, 0.f, 0.f);
        G.gVel2[id] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
      continue;
    }
    if (T.tmap[t] == ST_NONE) {  // no storage this substep: free fall, no grid
      for (uint32_t k = beg + threadIdx.x; k < end; k += AG2P_WG) adj_g2p_particle(P, T, G, T.binIdx[k], false, sCol, none, drop);
      continue;
    }
    const int3 org = make_int3(int32_t(t % P.tx) * 8 - 1, int32_t((t / P.tx) % P.ty) * 8 - 1, int32_t(t / (P.tx * P.ty)) * 8 - 1);
    for (uint32_t q = threadIdx.x; q < AWIN3; q += AG2P_WG) {
      const int32_t gi = org.x + int32_t(q % AWIN), gj = org.y + int32_t((q / AWIN) % AWIN), gk = org.z + int32_t(q / (AWIN * AWIN));
      const bool in = gi >= 0 && gj >= 0 && gk >= 0 && gi < gx && gj < gy && gk < gz;
      const uint32_t c = in ? V.cell(gi, gj, gk) * 3u : 0u;
      for (uint32_t a = 0; a < 3; a++) {
        sF[a][q] = in ? make_float2(__int_as_float(T.accVfin[c + a]), __int_as_float(T.accWfin[c + a])) : make_float2(0.f, 0.f);
        sN[a][q] = 0.f;
        sO[a][q] = 0.f;
      }
    }
    __syncthreads();
    // the waves interleave over each run of AG2P_WG bin entries (neighbouring
    // entries are mostly neighbouring particles, whose LDS adds would queue)
    static_assert(AG2P_WG == 512, "eight waves of 64 cover a run of AG2P_WG entries");
    for (uint32_t k0 = beg; k0 < end; k0 += AG2P_WG) {
      const uint32_t k = k0 + (threadIdx.x & 63u) * 8u + (threadIdx.x >> 6);
      if (k >= end) continue;
      adj_g2p_particle(
          P, T, G, T.binIdx[k], true, sCol,
          [&](uint32_t a, int32_t i, int32_t j, int32_t kk) { return sF[a][adj_window_face(P, a, i, j, kk, org)]; },
          [&](uint32_t a, int32_t i, int32_t j, int32_t kk, float dn, float dol) {
            const int32_t q = adj_window_face(P, a, i, j, kk, org);
            atomicAdd(&sN[a][q], dn);
            atomicAdd(&sO[a][q], dol);
          });
    }
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < AWIN3; q += AG2P_WG) {
      const int32_t gi = org.x + int32_t(q % AWIN), gj = org.y + int32_t((q / AWIN) % AWIN), gk = org.z + int32_t(q / (AWIN * AWIN));
      for (uint32_t a = 0; a < 3; a++) {
        const float dn = sN[a][q], dol = sO[a][q];
        if (dn == 0.f && dol == 0.f) continue;  // only faces particles sampled, all inside the grid
        const uint32_t f = V.cell(gi, gj, gk) * 3u + a;
        atomicAdd(&G.gFN[f], dn);
        atomicAdd(&G.gFO[f], dol);
      }
    }
    __syncthreads();
  }
  __syncthreads();
  if (threadIdx.x < min(P.ncols, uint32_t(ADJ_MAX_COLS)) * 3u && sCol[threadIdx.x] != 0.f)
    atomicAdd(&G.gCol[(threadIdx.x / 3u) * 9u + threadIdx.x % 3u], sCol[threadIdx.x]);
}

#define FOR_CELLS(T)                                                                            \
  for (uint32_t t = blockIdx.x * AWG + threadIdx.x, tot_ = (T).ctl[4] * 512u; t < tot_; t += gridDim.x * AWG)

// project (wgsl-solver.js:165-180) reversed: a face that was projected took
// -dt/h (p_hi - p_lo); fluid cells gather from their six faces.  The carry from
// the next substep's warm start is added here (fluid cells only: classify zeroes
// the others before the warm start).
__global__ void __launch_bounds__(AWG) k_adj_project(AdjTape T, AdjGrad G, uint32_t hasCarry) {
  const DParams& P = *T.P;
  const AView V{P, T};
  FOR_CELLS(T) {
    const uint32_t s = T.slotList[t >> 9], l = t & 511u, c = s * 512u + l;
    if ((T.ctype[c] & 0xffu) != 1u) {
      G.gPres[c] = 0.f;
      continue;
    }
    const int3 at = d_slotCoord(P, T.slotTile, s, l);
    const float k = P.dt * P.ih;
    float g = 0.f;
    for (uint32_t a = 0; a < 3; a++) {
      int3 lo = at, up = at;
      if (a == 0) { lo.x -= 1; up.x += 1; } else if (a == 1) { lo.y -= 1; up.y += 1; } else { lo.z -= 1; up.z += 1; }
      // own lower face: this cell is its high side
      const uint32_t tl = V.type(lo.x, lo.y, lo.z);
      if (tl != 2u) g -= k * G.gFN[c * 3u + a];
      // upper face (the lower face of the cell above): this cell is its low side
      const uint32_t tu = V.type(up.x, up.y, up.z);
      if (tu != 2u) g += k * G.gFN[V.cell(up.x, up.y, up.z) * 3u + a];
    }
    if (hasCarry) g += G.carry[T.slotTile[s] * 512u + l];
    G.gPres[c] = g;
  }
}

// The gather of a transposed sweep for every cell, fixed for the whole reversed
// solve (the topology is the forward pass's): the fluid neighbours a cell
// gathers from and their weights omega / n, in the order x-, x+, y-, y+, z-, z+.
__global__ void __launch_bounds__(AWG) k_adj_sor_prep(AdjTape T, AdjGrad G) {
  const DParams& P = *T.P;
  const AView V{P, T};
  const float w = P.omega;
  FOR_CELLS(T) {
    const uint32_t s = T.slotList[t >> 9], l = t & 511u, c = s * 512u + l;
    const bool fluid = (T.sten[c] & 0x80u) != 0u;
    const int3 at = d_slotCoord(P, T.slotTile, s, l);
    for (uint32_t q = 0; q < 6; q++) {
      uint32_t nb = ST_NONE;
      float coef = 0.f;
      if (fluid) {
        int3 r = at;
        if (q == 0) r.x -= 1; else if (q == 1) r.x += 1; else if (q == 2) r.y -= 1; else if (q == 3) r.y += 1; else if (q == 4) r.z -= 1; else r.z += 1;
        if (V.type(r.x, r.y, r.z) == 1u) {
          const uint32_t cm = V.cell(r.x, r.y, r.z);
          const uint32_t nm = T.sten[cm] & 7u;
          if (nm) {
            nb = cm;
            coef = w / float(nm);
          }
        }
      }
      G.sorNb[size_t(c) * 6u + q] = nb;
      G.sorCoef[size_t(c) * 6u + q] = coef;
    }
  }
}

// one colour of one red-black iteration, transposed: gin -> gout.  A cell's
// colour is that of its place in the tile (tile origins are even).
__device__ __forceinline__ void adj_sor_body(const AdjTape& T, const AdjGrad& G, const float* __restrict__ gin,
                                             float* __restrict__ gout, uint32_t colour, float w) {
  FOR_CELLS(T) {
    const uint32_t s = T.slotList[t >> 9], l = t & 511u, c = s * 512u + l;
    const uint32_t st = T.sten[c];
    if (!(st & 0x80u)) {
      gout[c] = 0.f;
      continue;
    }
    if ((((l & 7u) + ((l >> 3) & 7u) + (l >> 6)) & 1u) == colour) {
      const uint32_t n = st & 7u;
      if (n == 0u) {
        gout[c] = 0.f;
        continue;
      }
      const float g = gin[c];
      G.gRhs[c] += w / float(n) * g;
      gout[c] = (1.f - w) * g;
    } else {
      float g = gin[c];
      const uint32_t* nb = G.sorNb + size_t(c) * 6u;
      const float* cf = G.sorCoef + size_t(c) * 6u;
      for (uint32_t q = 0; q < 6; q++)
        if (nb[q] != ST_NONE) g += cf[q] * gin[nb[q]];
      gout[c] = g;
    }
  }
}
__global__ void __launch_bounds__(AWG) k_adj_sor(AdjTape T, AdjGrad G, const float* __restrict__ gin, float* __restrict__ gout,
                                                 uint32_t colour) {
  adj_sor_body(T, G, gin, gout, colour, T.P->omega);
}

// The whole transposed solve in one cooperative launch: the sweeps of
// k_adj_sor in the same order (per iteration black then red, ping-pong from
// gPres), a grid barrier between consecutive sweeps (the gathers of one read
// what the last wrote in other workgroups).  2 iters sweeps leave the result in
// gPres, as the launches per sweep do.  A barrier time-out (counted in
// bar[ADJ_BAR_FAILS]) ends the solve early; the host reports it.
__global__ void __launch_bounds__(AWG) k_adj_sor_all(AdjTape T, AdjGrad G) {
  __shared__ uint32_t sFail;
  if (threadIdx.x == 0) sFail = 0u;
  __syncthreads();
  const float w = T.P->omega;
  float* a = G.gPres;
  float* b = G.gPres2;
  uint32_t epoch = 0;
  for (uint32_t it = 0; it < T.iters; it++)
    for (int colour = 1; colour >= 0; colour--) {
      adj_sor_body(T, G, a, b, uint32_t(colour), w);
      if (it + 1u < T.iters || colour == 1)
        if (!gb_barrier(G.bar, G.bar + ADJ_BAR_RELEASE, G.bar + ADJ_BAR_FAILS, ++epoch, &sFail, blockIdx.x, gridDim.x)) return;
      float* t = a;
      a = b;
      b = t;
    }
}

__global__ void k_adj_clear_bar(uint32_t* __restrict__ bar) {
  for (uint32_t q = threadIdx.x; q < ADJ_BAR_WORDS; q += blockDim.x)
    if (q != ADJ_BAR_FAILS) bar[q] = 0u;
}

uint32_t adj_sor_coop_grid(uint32_t cus) {
  int dev = 0, coop = 0, nb = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess || !coop)
    return 0u;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_adj_sor_all, AWG, 0) != hipSuccess || nb < 1) return 0u;
  return std::min<uint32_t>(cus * uint32_t(nb), ADJ_SOR_MAX_WG);
}

// the gradient left on the warm start of this substep's solve, by tile of the
// box, for tiles that kept their slot from the previous substep (a new tile's
// pressure starts at zero)
__global__ void __launch_bounds__(AWG) k_adj_carry(AdjTape T, AdjGrad G, const float* __restrict__ g0) {
  FOR_CELLS(T) {
    const uint32_t s = T.slotList[t >> 9], l = t & 511u, tile = T.slotTile[s];
    if (T.tmap0[tile] == s) G.carry[tile * 512u + l] = g0[s * 512u + l];
  }
}

// divergence (106-123) reversed into the face gradient before projection, then
// normFaces (75-105) reversed into gradients of the sums V (gFN) and W (gFO); a
// face with a solid side held the collider's surface velocity v + w x (x - c)
// (new and old alike), whose gradient goes to that collider
__global__ void __launch_bounds__(AWG) k_adj_faces(AdjTape T, AdjGrad G) {
  const DParams& P = *T.P;
  const AView V{P, T};
  const float k = P.h * P.h / P.dt * P.ih;
  __shared__ float sCol[ADJ_MAX_COLS * 9];
  if (threadIdx.x < ADJ_MAX_COLS * 9) sCol[threadIdx.x] = 0.f;
  __syncthreads();
  const uint32_t ncol = min(P.ncols, uint32_t(ADJ_MAX_COLS));
  FOR_CELLS(T) {
    const uint32_t s = T.slotList[t >> 9], l = t & 511u, c = s * 512u + l;
    const int3 at = d_slotCoord(P, T.slotTile, s, l);
    const bool fluid = (T.ctype[c] & 0xffu) == 1u && d_inDom(P, at);
    const float gr = fluid ? G.gRhs[c] : 0.f;
    for (uint32_t a = 0; a < 3; a++) {
      int3 lo = at;
      if (a == 0) lo.x -= 1; else if (a == 1) lo.y -= 1; else lo.z -= 1;
      const uint32_t tl = V.type(lo.x, lo.y, lo.z), th = V.type(at.x, at.y, at.z);
      const uint32_t f = c * 3u + a;
      float gn = G.gFN[f] + k * gr;  // this face is the cell's lower face: div has -v/h
      if (tl == 1u && d_inDom(P, lo)) gn -= k * G.gRhs[V.cell(lo.x, lo.y, lo.z)];
      float gV = 0.f, gW = 0.f;
      const int32_t ws = T.accWraw[f];
      if (tl != 2u && th != 2u && ws > 0) {
        const float W = float(ws) * (1.0f / 65536.0f);
        const float gs = gn + G.gFO[f];
        gV = gs / W;
        gW = -gs * __int_as_float(T.accWfin[f]) / W;
      } else if ((tl == 2u || th == 2u) && ncol > 0u) {
        int32_t ci = -1;
        if (tl == 2u) ci = d_colAtCell(P, T.tmap, T.ctype, lo.x, lo.y, lo.z);
        if (ci < 0 && th == 2u) ci = d_colAtCell(P, T.tmap, T.ctype, at.x, at.y, at.z);
        const float gs = gn + G.gFO[f];
        if (ci >= 0 && ci < int32_t(ncol) && gs != 0.f) {
          // sv_a = v_a + (w x q)_a, q = x - c: d/dv = e_a, d/dw = q x e_a, d/dc = w x e_a
          const float3 fp = d_facePos(P, at, a);
          const float3 q = f3sub(fp, c3(T.C->c[ci * 8 + 5])), w = c3(T.C->c[ci * 8 + 7]);
          const float3 qe = a == 0 ? make_float3(0.f, q.z, -q.y) : (a == 1 ? make_float3(-q.z, 0.f, q.x) : make_float3(q.y, -q.x, 0.f));
          const float3 we = a == 0 ? make_float3(0.f, w.z, -w.y) : (a == 1 ? make_float3(-w.z, 0.f, w.x) : make_float3(w.y, -w.x, 0.f));
          float* o = sCol + ci * 9;
          atomicAdd(&o[0], gs * we.x); atomicAdd(&o[1], gs * we.y); atomicAdd(&o[2], gs * we.z);
          atomicAdd(&o[3 + a], gs);
          atomicAdd(&o[6], gs * qe.x); atomicAdd(&o[7], gs * qe.y); atomicAdd(&o[8], gs * qe.z);
        }
      }
      G.gFN[f] = gV;
      G.gFO[f] = gW;
    }
  }
  __syncthreads();
  if (threadIdx.x < ncol * 9u && sCol[threadIdx.x] != 0.f) atomicAdd(&G.gCol[threadIdx.x], sCol[threadIdx.x]);
}

// p2g (16-40) reversed: per particle, the 8 faces per axis it added to
__global__ void __launch_bounds__(AWG) k_adj_p2g(AdjTape T, AdjGrad G) {
  const DParams& P = *T.P;
  const AView V{P, T};
  const uint32_t i = blockIdx.x * AWG + threadIdx.x;
  if (i >= T.n) return;
  const float4 p4 = T.pos[i], v4 = T.vel[i];
  if (p4.y < ST_DEADY) return;
  const float3 p = c3(p4);
  const int3 fc = d_particleCell(P, p);
  if (T.tmap[d_tileOf(P, fc.x, fc.y, fc.z)] == ST_NONE) return;
  const uint32_t id = T.pid[i];
  const float va[3] = {v4.x, v4.y, v4.z};
  float gv[3] = {0.f, 0.f, 0.f};
  float3 gp = make_float3(0.f, 0.f, 0.f);
  for (uint32_t a = 0; a < 3; a++) {
    const float3 g = axisGrid(P, p, a);
    const float3 f0 = make_float3(floorf(g.x), floorf(g.y), floorf(g.z));
    const float3 fr = f3sub(g, f0);
    const int3 i0 = make_int3(int32_t(f0.x), int32_t(f0.y), int32_t(f0.z));
    for (uint32_t c = 0; c < 8; c++) {
      const int32_t dx = c & 1, dy = (c >> 1) & 1, dz = (c >> 2) & 1;
      const float wx = dx ? fr.x : 1.f - fr.x, wy = dy ? fr.y : 1.f - fr.y, wz = dz ? fr.z : 1.f - fr.z;
      const float sx = dx ? P.ih : -P.ih, sy = dy ? P.ih : -P.ih, sz = dz ? P.ih : -P.ih;
      const uint32_t f = V.face(a, i0.x + dx, i0.y + dy, i0.z + dz);
      const float gV = G.gFN[f], gW = G.gFO[f];
      gv[a] += wx * wy * wz * gV;
      const float s = gV * va[a] + gW;
      gp = f3add(gp, make_float3(sx * wy * wz * s, wx * sy * wz * s, wx * wy * sz * s));
    }
  }
  const float4 a = G.gPos2[id], b = G.gVel2[id];
  G.gPos2[id] = make_float4(a.x + gp.x, a.y + gp.y, a.z + gp.z, 0.f);
  G.gVel2[id] = make_float4(b.x + gv[0], b.y + gv[1], b.z + gv[2], 0.f);
}

__global__ void k_adj_zero(float* __restrict__ p, size_t n) {
  for (size_t i = size_t(blockIdx.x) * AWG + threadIdx.x; i < n; i += size_t(gridDim.x) * AWG) p[i] = 0.f;
}

static uint32_t grid_for(size_t n) { return uint32_t(std::min<size_t>((n + AWG - 1) / AWG, 8192)); }

bool launch_adjoint_step(hipStream_t st, const AdjTape& T, const AdjGrad& G, uint32_t scap, uint32_t ntiles, uint32_t coopGrid) {
  const size_t C = size_t(scap) * 512u;
  const uint32_t cg = std::max<uint32_t>(1u, std::min<uint32_t>(scap * 2u, 4096u));
  const uint32_t pg = uint32_t((T.n + AWG - 1) / AWG);
  hipLaunchKernelGGL(k_adj_zero, dim3(grid_for(C * 3)), dim3(AWG), 0, st, G.gFN, C * 3);
  hipLaunchKernelGGL(k_adj_zero, dim3(grid_for(C * 3)), dim3(AWG), 0, st, G.gFO, C * 3);
  hipLaunchKernelGGL(k_adj_zero, dim3(grid_for(C)), dim3(AWG), 0, st, G.gRhs, C);
  hipLaunchKernelGGL(k_adj_zero, dim3(1), dim3(AWG), 0, st, G.gCol, size_t(ADJ_MAX_COLS * 9));
  if (pg) hipLaunchKernelGGL(k_adj_g2p, dim3(std::min<uint32_t>(ntiles + 1u, 8192u)), dim3(AG2P_WG), 0, st, T, G);
  hipLaunchKernelGGL(k_adj_project, dim3(cg), dim3(AWG), 0, st, T, G, 1u);
  // the solve reversed: per iteration, black then red, ping-pong gPres <-> gPres2
  float* a = G.gPres;
  float* b = G.gPres2;
  hipLaunchKernelGGL(k_adj_sor_prep, dim3(cg), dim3(AWG), 0, st, T, G);
  bool one = false;
  if (coopGrid && T.iters) {
    hipLaunchKernelGGL(k_adj_clear_bar, dim3(1), dim3(AWG), 0, st, G.bar);
    AdjTape t = T;
    AdjGrad g = G;
    void* args[] = {(void*)&t, (void*)&g};
    one = hipLaunchCooperativeKernel((const void*)k_adj_sor_all, dim3(std::min(coopGrid, ADJ_SOR_MAX_WG)), dim3(AWG), args, 0, st) == hipSuccess;
    if (!one) (void)hipGetLastError();  // the launches per sweep below do the solve
  }
  for (uint32_t it = 0; !one && it < T.iters; it++)
    for (int colour = 1; colour >= 0; colour--) {
      hipLaunchKernelGGL(k_adj_sor, dim3(cg), dim3(AWG), 0, st, T, G, (const float*)a, b, uint32_t(colour));
      std::swap(a, b);
    }
  // a holds the gradient of the warm start: the carry for the substep before
  hipLaunchKernelGGL(k_adj_zero, dim3(grid_for(size_t(ntiles) * 512u)), dim3(AWG), 0, st, G.carry, size_t(ntiles) * 512u);
  hipLaunchKernelGGL(k_adj_carry, dim3(cg), dim3(AWG), 0, st, T, G, (const float*)a);
  hipLaunchKernelGGL(k_adj_faces, dim3(cg), dim3(AWG), 0, st, T, G);
  if (pg) hipLaunchKernelGGL(k_adj_p2g, dim3(pg), dim3(AWG), 0, st, T, G);
  return one;
}

// flip_norms.hip - FP64 norms of the pressure solve and of the projected face
// velocities: the only f64 arithmetic on the GPU.
//
// After a substep the pool still holds what its pressure stage solved and what
// its projection produced: cell types, particle counts, the right-hand side,
// the pressure, and the projected face velocities (the reference's `accV`
// words as f32 bits, read by g2p and left in place).  For every fluid cell
// (type 1) this pass forms, in f64 from those f32 values:
//   r   = rhs + sum of fluid-neighbour pressures - nns * p   (the residual of
//         the reference's SOR system, relax at src/flip/wgsl-solver.js:124-140;
//         nns = non-solid neighbours, a cell with nns = 0 counts as r = 0)
//   div = (sum over axes of v_hi - v_lo) / h                 (divergence at
//         wgsl-solver.js:106-113, of the projected faces)
//   e   = div - tgt, tgt the density-correction target of the same cell
//         (wgsl-solver.js:114-121), i.e. what the solve drives div to.
// In exact arithmetic the projection (wgsl-solver.js:165-180) gives
// e = -(dt / h^2) r, so e measures what the solve left undone.
//
// Reduction: each thread sums its cells in order, a wave combines its lanes in
// a fixed butterfly, a workgroup its four waves in order, and one workgroup
// combines the per-workgroup partials in a fixed tree: the result does not
// depend on scheduling.
#include <mutex>

#include "flip_device.h"
#include "flip_kernels.h"

#define NWG_T 256

namespace {

__device__ __forceinline__ double wave_sum_f64(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
__device__ __forceinline__ double wave_max_f64(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  return v;
}

// pressure of a neighbour cell known to be fluid (inside the box)
__device__ __forceinline__ double pres_at(const DParams& P, const DGrid& G, int32_t i, int32_t j, int32_t k) {
  return double(G.pres[d_cellIdx(P, G.tmap, i, j, k)]);
}
__device__ __forceinline__ double face_v(const DParams& P, const DGrid& G, uint32_t axis, int32_t i, int32_t j, int32_t k) {
  return double(st_u2f(uint32_t(G.accV[d_faceIndex(P, G.tmap, axis, i, j, k)])));
}

}  // namespace

__global__ void __launch_bounds__(NWG_T) k_norms(const DParams* __restrict__ PP, DGrid G, double* __restrict__ part) {
  const DParams& P = *PP;
  double acc[ST_NORM_WORDS];
  for (int q = 0; q < ST_NORM_WORDS; q++) acc[q] = 0.0;
  const double h = double(P.h), rest = double(P.rest), comp = double(P.comp);
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, tot = G.ctl[4] * 512u; t < tot; t += gridDim.x * blockDim.x) {
    uint32_t s = G.slotList[t >> 9], local = t & 511u, c = s * 512u + local;
    uint32_t ct = G.ctype[c];
    if ((ct & 0xffu) != 1u) continue;
    int3 a = d_slotCoord(P, G.slotTile, s, local);
    int32_t i = a.x, j = a.y, k = a.z;
    const int32_t di[6] = {-1, 1, 0, 0, 0, 0}, dj[6] = {0, 0, -1, 1, 0, 0}, dk[6] = {0, 0, 0, 0, -1, 1};
    double sum = 0.0;
    int nns = 0;
    bool interior = true;
    for (int q = 0; q < 6; q++) {
      uint32_t ty = d_ctypeOf(P, G.tmap, G.ctype, i + di[q], j + dj[q], k + dk[q]);
      if (ty == 0u) interior = false;
      if (ty != 2u) nns++;
      if (ty == 1u) sum += pres_at(P, G, i + di[q], j + dj[q], k + dk[q]);
    }
    double rhs = double(G.rhs[c]);
    double r = nns == 0 ? 0.0 : rhs + sum - double(nns) * double(G.pres[c]);
    double div = ((face_v(P, G, 0, i + 1, j, k) - face_v(P, G, 0, i, j, k)) + (face_v(P, G, 1, i, j + 1, k) - face_v(P, G, 1, i, j, k)) +
                  (face_v(P, G, 2, i, j, k + 1) - face_v(P, G, 2, i, j, k))) / h;
    double dens = double(G.cnt[c]) / rest - 1.0;
    double tgt = comp * (interior ? dens : fmax(dens, 0.0));
    if (P.nb != 0u && d_phiOfBits(ct) < -1.0f) tgt = 0.0;
    double e = div - tgt;
    acc[0] += r * r;
    acc[1] = fmax(acc[1], fabs(r));
    acc[2] += rhs * rhs;
    acc[3] += div * div;
    acc[4] = fmax(acc[4], fabs(div));
    acc[5] += e * e;
    acc[6] = fmax(acc[6], fabs(e));
    acc[7] += 1.0;
  }
  __shared__ double red[NWG_T / 64][ST_NORM_WORDS];
  uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  for (int q = 0; q < ST_NORM_WORDS; q++) {
    bool isMax = q == 1 || q == 4 || q == 6;
    double v = isMax ? wave_max_f64(acc[q]) : wave_sum_f64(acc[q]);
    if (lane == 0) red[w][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < ST_NORM_WORDS) {
    int q = int(threadIdx.x);
    bool isMax = q == 1 || q == 4 || q == 6;
    double v = red[0][q];
    for (uint32_t u = 1; u < NWG_T / 64; u++) v = isMax ? fmax(v, red[u][q]) : v + red[u][q];
    part[size_t(blockIdx.x) * ST_NORM_WORDS + q] = v;
  }
}

// one workgroup: partials of nwg workgroups -> out[ST_NORM_WORDS]
__global__ void __launch_bounds__(NWG_T) k_norms_final(const double* __restrict__ part, uint32_t nwg, double* __restrict__ out) {
  __shared__ double red[NWG_T];
  for (int q = 0; q < ST_NORM_WORDS; q++) {
    bool isMax = q == 1 || q == 4 || q == 6;
    double v = 0.0;
    for (uint32_t b = threadIdx.x; b < nwg; b += NWG_T) {
      double x = part[size_t(b) * ST_NORM_WORDS + q];
      v = isMax ? fmax(v, x) : v + x;
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t half = NWG_T / 2; half > 0; half >>= 1) {
      if (threadIdx.x < half) red[threadIdx.x] = isMax ? fmax(red[threadIdx.x], red[threadIdx.x + half]) : red[threadIdx.x] + red[threadIdx.x + half];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = red[0];
    __syncthreads();
  }
}

void launch_norms(hipStream_t st, const LaunchCtx& L, double* part, double* out) {
  hipLaunchKernelGGL(k_norms, dim3(L.cellGrid), dim3(NWG_T), 0, st, L.P, L.G, part);
  hipLaunchKernelGGL(k_norms_final, dim3(1), dim3(NWG_T), 0, st, part, L.cellGrid, out);
}

// The allocation-free norms (st_solve_norms_static): the per-workgroup partials
// and the results live in this fixed array, one copy per GPU for the life of the
// process.  In host code `scratch` names a host-side stand-in, so its device
// address on each GPU is looked up once (with that GPU current) and kept.
static __device__ __align__(256) double scratch[(NORMS_STATIC_WG + 1) * ST_NORM_WORDS];

namespace {
std::mutex g_scratchMu;
double* g_scratchDev[NORMS_STATIC_DEVICES] = {};
}  // namespace

hipError_t norms_static_scratch(int dev, double** ptr, size_t* words) {
  if (dev < 0 || dev >= NORMS_STATIC_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratchMu);
  if (!g_scratchDev[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratchDev[dev] = static_cast<double*>(p);
  }
  *ptr = g_scratchDev[dev];
  *words = size_t(NORMS_STATIC_WG + 1) * ST_NORM_WORDS;
  return hipSuccess;
}

// flip_sdf.hip - the collider distance-grid builder (host code).
//
// Restates buildSDF (src/shared.js:98-181), the reference's signed distance
// grid of a closed triangle mesh, which it runs on the CPU in JS doubles and
// stores in a Float32Array: exact squared distances near the surface (each
// triangle visits the cells within 2.5 cells of it), square roots, two
// chamfer sweeps over the 13 already-passed neighbours (forward, then
// backward), then the sign from crossings of lines along x.  Every quantity is
// computed in f64 in the reference's order and stored through f32 where the
// reference stores through its Float32Array, so the grid is the reference's
// bit for bit.  The sweeps are sequential by definition (each cell reads the
// values the sweep has just written); a grid is built once per mesh edit.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "splashtank.h"

namespace {

// Math.hypot as JS engines evaluate it for three finite arguments: scaled by the
// largest magnitude, squares summed with Kahan compensation, root, rescale
double js_hypot3(double a, double b, double c) {
  double m = std::max(std::fabs(a), std::max(std::fabs(b), std::fabs(c)));
  if (m == 0.0) return 0.0;
  double sum = 0.0, comp = 0.0;
  const double xs[3] = {a, b, c};
  for (double x : xs) {
    double n = x / m;
    double summand = n * n - comp;
    double prelim = sum + summand;
    comp = (prelim - sum) - summand;
    sum = prelim;
  }
  return std::sqrt(sum) * m;
}

// ptTriDist2 (src/shared.js:68-96): squared distance from p to triangle abc (Ericson)
double pt_tri_dist2(double px, double py, double pz, const double* a, const double* b, const double* c) {
  const double abx = b[0] - a[0], aby = b[1] - a[1], abz = b[2] - a[2], acx = c[0] - a[0], acy = c[1] - a[1], acz = c[2] - a[2];
  const double apx = px - a[0], apy = py - a[1], apz = pz - a[2];
  const double d1 = abx * apx + aby * apy + abz * apz, d2 = acx * apx + acy * apy + acz * apz;
  double qx, qy, qz;
  if (d1 <= 0 && d2 <= 0) {
    qx = a[0]; qy = a[1]; qz = a[2];
  } else {
    const double bpx = px - b[0], bpy = py - b[1], bpz = pz - b[2];
    const double d3 = abx * bpx + aby * bpy + abz * bpz, d4 = acx * bpx + acy * bpy + acz * bpz;
    if (d3 >= 0 && d4 <= d3) {
      qx = b[0]; qy = b[1]; qz = b[2];
    } else {
      const double vc = d1 * d4 - d3 * d2;
      if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        const double v = d1 / (d1 - d3);
        qx = a[0] + abx * v; qy = a[1] + aby * v; qz = a[2] + abz * v;
      } else {
        const double cpx = px - c[0], cpy = py - c[1], cpz = pz - c[2];
        const double d5 = abx * cpx + aby * cpy + abz * cpz, d6 = acx * cpx + acy * cpy + acz * cpz;
        if (d6 >= 0 && d5 <= d6) {
          qx = c[0]; qy = c[1]; qz = c[2];
        } else {
          const double vb = d5 * d2 - d1 * d6;
          if (vb <= 0 && d2 >= 0 && d6 <= 0) {
            const double w = d2 / (d2 - d6);
            qx = a[0] + acx * w; qy = a[1] + acy * w; qz = a[2] + acz * w;
          } else {
            const double va = d3 * d6 - d5 * d4;
            if (va <= 0 && d4 - d3 >= 0 && d5 - d6 >= 0) {
              const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
              qx = b[0] + (c[0] - b[0]) * w; qy = b[1] + (c[1] - b[1]) * w; qz = b[2] + (c[2] - b[2]) * w;
            } else {
              const double den = 1 / (va + vb + vc), v = vb * den, w = vc * den;
              qx = a[0] + abx * v + acx * w; qy = a[1] + aby * v + acy * w; qz = a[2] + abz * v + acz * w;
            }
          }
        }
      }
    }
  }
  const double dx = px - qx, dy = py - qy, dz = pz - qz;
  return dx * dx + dy * dy + dz * dz;
}

}  // namespace

extern "C" {

int st_build_sdf(const double* V, uint32_t nv, const uint32_t* tri, uint32_t nt, double cell, double margin, st_sdf_grid* grid,
                 float* data, size_t cap) {
  if (!V || !tri || !grid || nv == 0 || nt == 0) return ST_ERR_ARG;
  for (size_t q = 0; q < size_t(nt) * 3; q++)
    if (tri[q] >= nv) return ST_ERR_ARG;
  if (!(cell > 0)) cell = 0.025;
  if (!(margin >= 0)) margin = 0.16;
  // bounds (meshBounds, src/shared.js:63-67) widened by the margin
  double lo[3] = {1e9, 1e9, 1e9}, hi[3] = {-1e9, -1e9, -1e9};
  for (uint32_t p = 0; p < nv; p++)
    for (int k = 0; k < 3; k++) {
      if (V[p * 3 + k] < lo[k]) lo[k] = V[p * 3 + k];
      if (V[p * 3 + k] > hi[k]) hi[k] = V[p * 3 + k];
    }
  for (int k = 0; k < 3; k++) { lo[k] = lo[k] - margin; hi[k] = hi[k] + margin; }
  double dims[3];
  auto setDims = [&]() { for (int k = 0; k < 3; k++) dims[k] = std::ceil((hi[k] - lo[k]) / cell) + 1; };
  setDims();
  while (dims[0] * dims[1] * dims[2] > 2.6e6) {
    cell *= 1.25;
    setDims();
  }
  const int32_t nx = int32_t(dims[0]), ny = int32_t(dims[1]), nz = int32_t(dims[2]);
  const size_t nxy = size_t(nx) * ny, N = nxy * nz;
  for (int k = 0; k < 3; k++) { grid->dims[k] = int32_t(dims[k]); grid->origin[k] = lo[k]; }
  grid->cell = cell;
  if (!data) return ST_OK;
  if (cap < N) return ST_ERR_ARG;
  const double ic = 1 / cell;
  std::fill(data, data + N, float(1e30));
  // 1. exact squared distances in a band of 2.5 cells around every triangle
  const double B = 2.5 * cell, B2 = B * B;
  for (uint32_t q = 0; q < nt; q++) {
    const double *a = V + size_t(tri[q * 3]) * 3, *b = V + size_t(tri[q * 3 + 1]) * 3, *c = V + size_t(tri[q * 3 + 2]) * 3;
    const double ux = b[0] - a[0], uy = b[1] - a[1], uz = b[2] - a[2], vx = c[0] - a[0], vy = c[1] - a[1], vz = c[2] - a[2];
    double px = uy * vz - uz * vy, py = uz * vx - ux * vz, pz = ux * vy - uy * vx;
    const double pl = js_hypot3(px, py, pz);
    if (pl > 0) { px /= pl; py /= pl; pz /= pl; }
    auto rlo = [&](int k) {
      double f = std::floor((std::min(a[k], std::min(b[k], c[k])) + (-B) - lo[k]) * ic);
      return std::max(0.0, std::min(dims[k] - 1, f));
    };
    auto rhi = [&](int k) {
      double f = std::ceil((std::max(a[k], std::max(b[k], c[k])) + B - lo[k]) * ic);
      return std::max(0.0, std::min(dims[k] - 1, f));
    };
    const int32_t i0 = int32_t(rlo(0)), i1 = int32_t(rhi(0)), j0 = int32_t(rlo(1)), j1 = int32_t(rhi(1)), k0 = int32_t(rlo(2)),
                  k1 = int32_t(rhi(2));
    for (int32_t k = k0; k <= k1; k++) {
      const double z = lo[2] + k * cell;
      for (int32_t j = j0; j <= j1; j++) {
        const double y = lo[1] + j * cell;
        const size_t base = size_t(j) * nx + size_t(k) * nxy;
        for (int32_t i = i0; i <= i1; i++) {
          const double x = lo[0] + i * cell;
          if (pl > 0) {
            const double e = (x - a[0]) * px + (y - a[1]) * py + (z - a[2]) * pz;
            if (e * e > B2) continue;
          }
          const double d2 = pt_tri_dist2(x, y, z, a, b, c);
          if (d2 < double(data[base + i])) data[base + i] = float(d2);
        }
      }
    }
  }
  for (size_t i = 0; i < N; i++)
    if (double(data[i]) < 1e29) data[i] = float(std::sqrt(double(data[i])));
  // 2. chamfer sweeps: forward over the 13 neighbours already passed, backward over the other 13
  int32_t oi[13], oj[13], ok[13];
  int64_t od[13];
  float ow[13];
  {
    int o = 0;
    for (int dk = -1; dk <= 0; dk++)
      for (int dj = -1; dj <= 1; dj++)
        for (int di = -1; di <= 1; di++)
          if (dk < 0 || dj < 0 || (dj == 0 && di < 0)) {
            oi[o] = di; oj[o] = dj; ok[o] = dk;
            od[o] = int64_t(di) + int64_t(dj) * nx + int64_t(dk) * int64_t(nxy);
            ow[o] = float(cell * js_hypot3(di, dj, dk));
            o++;
          }
  }
  for (int pass = 0; pass < 2; pass++) {
    const int64_t sg = pass ? -1 : 1;
    for (int32_t kk = 0; kk < nz; kk++) {
      const int32_t k = pass ? nz - 1 - kk : kk;
      const bool kIn = k > 0 && k < nz - 1;
      for (int32_t jj = 0; jj < ny; jj++) {
        const int32_t j = pass ? ny - 1 - jj : jj;
        const bool jkIn = kIn && j > 0 && j < ny - 1;
        const int64_t rowBase = int64_t(j) * nx + int64_t(k) * int64_t(nxy);
        for (int32_t ii = 0; ii < nx; ii++) {
          const int32_t i = pass ? nx - 1 - ii : ii;
          const int64_t idx = rowBase + i;
          double v = data[idx];
          if (v <= B) continue;  // exact already
          if (jkIn && i > 0 && i < nx - 1) {
            for (int o = 0; o < 13; o++) {
              const double c2 = double(data[idx + sg * od[o]]) + double(ow[o]);
              if (c2 < v) v = c2;
            }
          } else {
            for (int o = 0; o < 13; o++) {
              const int64_t ni = i + sg * oi[o], nj = j + sg * oj[o], nk = k + sg * ok[o];
              if (ni < 0 || nj < 0 || nk < 0 || ni >= nx || nj >= ny || nk >= nz) continue;
              const double c2 = double(data[idx + sg * od[o]]) + double(ow[o]);
              if (c2 < v) v = c2;
            }
          }
          data[idx] = float(v);
        }
      }
    }
  }
  // 3. inside or outside: crossings of each line along x, nudged off the grid
  const double ey = 0.000137 * cell, ez = 0.000271 * cell;
  std::vector<std::vector<double>> rows(size_t(ny) * nz);
  for (uint32_t q = 0; q < nt; q++) {
    const double *a = V + size_t(tri[q * 3]) * 3, *b = V + size_t(tri[q * 3 + 1]) * 3, *c = V + size_t(tri[q * 3 + 2]) * 3;
    const double area = (b[1] - a[1]) * (c[2] - a[2]) - (b[2] - a[2]) * (c[1] - a[1]);
    if (std::fabs(area) < 1e-14) continue;  // edge-on to the lines
    const double jlo = std::max(0.0, std::ceil((std::min(a[1], std::min(b[1], c[1])) - lo[1] - ey) * ic));
    const double jhi = std::min(double(ny - 1), std::floor((std::max(a[1], std::max(b[1], c[1])) - lo[1] - ey) * ic));
    const double klo = std::max(0.0, std::ceil((std::min(a[2], std::min(b[2], c[2])) - lo[2] - ez) * ic));
    const double khi = std::min(double(nz - 1), std::floor((std::max(a[2], std::max(b[2], c[2])) - lo[2] - ez) * ic));
    for (double kd = klo; kd <= khi; kd++) {
      const int32_t k = int32_t(kd);
      const double z = lo[2] + k * cell + ez;
      for (double jd = jlo; jd <= jhi; jd++) {
        const int32_t j = int32_t(jd);
        const double y = lo[1] + j * cell + ey;
        const double w0 = ((c[1] - b[1]) * (z - b[2]) - (c[2] - b[2]) * (y - b[1])) / area;
        const double w1 = ((a[1] - c[1]) * (z - c[2]) - (a[2] - c[2]) * (y - c[1])) / area;
        const double w2 = 1 - w0 - w1;
        if (w0 < 0 || w1 < 0 || w2 < 0) continue;
        rows[size_t(j) + size_t(ny) * k].push_back(w0 * a[0] + w1 * b[0] + w2 * c[0]);
      }
    }
  }
  for (int32_t k = 0; k < nz; k++)
    for (int32_t j = 0; j < ny; j++) {
      std::vector<double>& xs = rows[size_t(j) + size_t(ny) * k];
      if (xs.empty()) continue;
      std::sort(xs.begin(), xs.end());
      size_t m = 0;
      const size_t base = size_t(j) * nx + size_t(k) * nxy;
      for (int32_t i = 0; i < nx; i++) {
        const double x = lo[0] + i * cell;
        while (m < xs.size() && xs[m] < x) m++;
        if (m & 1) data[base + i] = -data[base + i];
      }
    }
  return ST_OK;
}

}  // extern "C"

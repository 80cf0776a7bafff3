'use strict';
// The collider distance grid, as the reference's buildSDF(mesh, cell, margin)
// (src/shared.js:98-181) returns it: {id, dims, origin, cell, data}, built by
// st_build_sdf in the native library (host code, the reference's arithmetic
// bit for bit).  Meshes are {v: [[x, y, z], ...], f: [[i0, i1, i2, ...], ...]}
// with polygon faces, fan-triangulated as the reference's meshTris does
// (src/shared.js:46-50).
const path = require('path');

let addon = null;
function loadAddon() {
  if (!addon) addon = require(path.join(__dirname, 'splashtank_hip.node'));
  return addon;
}
let SDF_ID = 1;
function meshTris(m) {
  const t = [], of = [];
  m.f.forEach((face, fi) => { for (let k = 1; k + 1 < face.length; k++) { t.push(face[0], face[k], face[k + 1]); of.push(fi); } });
  return { t, of };
}
function buildSDF(mesh, cell, margin) {
  cell = cell || 0.025; margin = margin === undefined ? 0.16 : margin;
  const v = new Float64Array(mesh.v.length * 3);
  mesh.v.forEach((p, i) => { v[i * 3] = p[0]; v[i * 3 + 1] = p[1]; v[i * 3 + 2] = p[2]; });
  const g = loadAddon().buildSdf(v, Uint32Array.from(meshTris(mesh).t), cell, margin);
  return { id: SDF_ID++, dims: g.dims, origin: g.origin, cell: g.cell, data: g.data };
}
// drip samples on a collider, from its distance grid (dripSamples, src/flip/bake.js:17-40): grid points within
// 0.6 cells of the surface moved onto it along the gradient, about one per solver cell of surface (100 to 6000),
// picked by a seeded stride; 8 floats each: local position, collider number k, local normal, wetness 0 (dry)
const DRIP_CACHE = new Map();
function dripSamples(sdf, h, k) {
  const key = sdf.id + '|' + h.toFixed(5);
  let base = DRIP_CACHE.get(key);
  if (!base) {
    const nx = sdf.dims[0], ny = sdf.dims[1], nz = sdf.dims[2], c = sdf.cell, d = sdf.data, o = sdf.origin, pts = [];
    const at = (i, j, q) => d[Math.min(nx - 1, Math.max(0, i)) + nx * (Math.min(ny - 1, Math.max(0, j)) + ny * Math.min(nz - 1, Math.max(0, q)))];
    for (let q = 1; q < nz - 1; q++) {
      for (let j = 1; j < ny - 1; j++) {
        for (let i = 1; i < nx - 1; i++) {
          const v = d[i + nx * (j + ny * q)];
          if (Math.abs(v) > 0.6 * c) continue;
          let gx = at(i + 1, j, q) - at(i - 1, j, q), gy = at(i, j + 1, q) - at(i, j - 1, q), gz = at(i, j, q + 1) - at(i, j, q - 1);
          const gl = Math.hypot(gx, gy, gz);
          if (gl < 1e-9) continue;
          gx /= gl; gy /= gl; gz /= gl;
          pts.push([o[0] + i * c - gx * v, o[1] + j * c - gy * v, o[2] + q * c - gz * v, gx, gy, gz]);
        }
      }
    }
    const nsurf = pts.length, want = Math.max(100, Math.min(6000, Math.round(nsurf * c * c / (h * h))));
    const step = Math.max(1, nsurf / want);
    let sd = 7 + (sdf.id | 0);
    sd = (sd * 16807) % 2147483647;
    base = [];
    for (let t = (sd / 2147483647) * step; t < nsurf; t += step) { const p = pts[Math.floor(t)]; base.push(p[0], p[1], p[2], 0, p[3], p[4], p[5], 0); }
    base = new Float32Array(base);
    DRIP_CACHE.set(key, base);
    if (DRIP_CACHE.size > 32) DRIP_CACHE.delete(DRIP_CACHE.keys().next().value);
  }
  const a = base.slice();
  for (let i = 3; i < a.length; i += 8) a[i] = k;
  return a;
}
module.exports = { buildSDF, meshTris, dripSamples };

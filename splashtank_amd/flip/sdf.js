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
// Drip samples on a collider (dripSamples, src/flip/bake.js:17-40): the grid
// nodes of its distance field within 0.6 cells of the surface, each moved onto
// the surface along the field's normalised central-difference gradient (indices
// clamped at the grid's edge); of those, about one per solver cell of surface
// (100 to 6000) taken at an even stride from a start drawn by the Park-Miller
// minimal standard generator (16807 mod 2^31 - 1) seeded 7 + sdf.id.  8 floats
// each: local position, collider number k, local normal, wetness 0 (dry).  The
// picks depend only on the grid and h, so they are kept per (grid, h), 32 at most.
function surfacePoints(sdf) {
  const [nx, ny, nz] = sdf.dims, cell = sdf.cell, D = sdf.data, org = sdf.origin;
  const clampIdx = (v, n) => Math.min(n - 1, Math.max(0, v));
  const value = (i, j, q) => D[clampIdx(i, nx) + nx * (clampIdx(j, ny) + ny * clampIdx(q, nz))];
  const pts = [];
  for (let q = 1; q < nz - 1; q++) {
    for (let j = 1; j < ny - 1; j++) {
      for (let i = 1; i < nx - 1; i++) {
        const dist = D[i + nx * (j + ny * q)];
        if (Math.abs(dist) > 0.6 * cell) continue;
        const g = [value(i + 1, j, q) - value(i - 1, j, q), value(i, j + 1, q) - value(i, j - 1, q), value(i, j, q + 1) - value(i, j, q - 1)];
        const len = Math.hypot(g[0], g[1], g[2]);
        if (len < 1e-9) continue;
        const n = g.map((x) => x / len);
        pts.push({ p: [org[0] + i * cell - n[0] * dist, org[1] + j * cell - n[1] * dist, org[2] + q * cell - n[2] * dist], n });
      }
    }
  }
  return pts;
}
const DRIP_PICKS = new Map();
function dripSamples(sdf, h, k) {
  const key = sdf.id + '|' + h.toFixed(5);
  let picks = DRIP_PICKS.get(key);
  if (!picks) {
    const pts = surfacePoints(sdf), cell = sdf.cell;
    const want = Math.max(100, Math.min(6000, Math.round(pts.length * cell * cell / (h * h))));
    const stride = Math.max(1, pts.length / want);
    const seed = (((7 + (sdf.id | 0)) * 16807) % 2147483647) / 2147483647;
    const rows = [];
    for (let t = seed * stride; t < pts.length; t += stride) {
      const s = pts[Math.floor(t)];
      rows.push(s.p[0], s.p[1], s.p[2], 0, s.n[0], s.n[1], s.n[2], 0);
    }
    picks = new Float32Array(rows);
    DRIP_PICKS.set(key, picks);
    if (DRIP_PICKS.size > 32) DRIP_PICKS.delete(DRIP_PICKS.keys().next().value);
  }
  const out = picks.slice();
  for (let r = 0; r < out.length / 8; r++) out[r * 8 + 3] = k;
  return out;
}
module.exports = { buildSDF, meshTris, dripSamples };

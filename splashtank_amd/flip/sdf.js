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
module.exports = { buildSDF, meshTris };

'use strict';
// Writes the golden vectors of the reference's collider SDF builder
// (buildSDF, src/shared.js:98-181) for three meshes the reference makes itself
// (meshCube, meshSphere, meshCylinder, src/shared.js:15-45), at the settings its
// page uses for colliders (cell 0.025, margin 0.16, src/app/03-objects.js:16).
// Run once in a container that holds the reference:
//   node tests/golden/make_sdf_golden.js /root/reference/src/shared.js tests/golden
// Output: sdf_golden.json (meshes as vertex and triangle lists, grid dims, origin,
// cell) and sdf_golden.bin (the f32 grids back to back, little endian).
const fs = require('fs');
const vm = require('vm');
const path = require('path');
const [src, outDir] = process.argv.slice(2);
const ctx = vm.createContext({ Math, Float32Array, Float64Array, Int32Array, Uint32Array, Array, Map, Object, JSON, Number, console });
vm.runInContext(fs.readFileSync(src, 'utf8') + '\n;this.__api = { meshCube, meshSphere, meshCylinder, meshTris, buildSDF };', ctx);
const api = ctx.__api;
const cases = [
  { name: 'cube', mesh: api.meshCube(0.1, 0.1, 0.1) },
  { name: 'sphere', mesh: api.meshSphere(0.12, 12, 8) },
  { name: 'cylinder', mesh: api.meshCylinder(0.08, 0.15, 10) },
];
const meta = [];
const blobs = [];
let off = 0;
for (const c of cases) {
  const sd = api.buildSDF(c.mesh, 0.025, 0.16);
  const tris = Array.from(api.meshTris(c.mesh).t);
  meta.push({ name: c.name, v: c.mesh.v.map((p) => p.slice()), tris, cell: 0.025, margin: 0.16,
    dims: sd.dims.slice(), origin: sd.origin.slice(), sdfCell: sd.cell, offset: off, count: sd.data.length });
  blobs.push(Buffer.from(sd.data.buffer, sd.data.byteOffset, sd.data.byteLength));
  off += sd.data.length;
}
fs.writeFileSync(path.join(outDir, 'sdf_golden.json'), JSON.stringify({ source: 'buildSDF (src/shared.js:98-181)', cases: meta }));
fs.writeFileSync(path.join(outDir, 'sdf_golden.bin'), Buffer.concat(blobs));
console.log(meta.map((m) => `${m.name}: dims ${m.dims.join('x')}, ${m.tris.length / 3} triangles`).join('\n'));

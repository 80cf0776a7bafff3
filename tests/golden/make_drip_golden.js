'use strict';
// Writes the golden vectors of the reference's drip samples (dripSamples,
// src/flip/bake.js:19-40): points over a collider's surface taken from its
// distance grid.  The grids are the reference's own buildSDF grids of
// tests/golden/sdf_golden.* (ids 1..3 as the reference numbers them), sampled
// for solver cells h = 1/64 and 1/32, collider number k = 2.  Run once in a
// container that holds the reference:
//   node tests/golden/make_drip_golden.js /root/reference/src/flip/bake.js tests/golden
// Output: drip_golden.json (cases: grid name, id, h, k, sample count, offset)
// and drip_golden.bin (8 f32 per sample, little endian).
const fs = require('fs');
const vm = require('vm');
const path = require('path');
const [src, dir] = process.argv.slice(2);
const ctx = vm.createContext({ Math, Float32Array, Float64Array, Int32Array, Uint32Array, Array, Map, Object, JSON, Number, console });
vm.runInContext(fs.readFileSync(src, 'utf8') + '\n;this.__api = { dripSamples };', ctx);
const meta = JSON.parse(fs.readFileSync(path.join(dir, 'sdf_golden.json'), 'utf8'));
const raw = fs.readFileSync(path.join(dir, 'sdf_golden.bin'));
const all = new Float32Array(raw.buffer.slice(raw.byteOffset, raw.byteOffset + raw.byteLength));
const cases = [];
const blobs = [];
let off = 0;
meta.cases.forEach((c, q) => {
  const sdf = { id: q + 1, dims: c.dims, origin: c.origin, cell: c.sdfCell, data: all.subarray(c.offset, c.offset + c.count) };
  for (const h of [1 / 64, 1 / 32]) {
    const a = ctx.__api.dripSamples(sdf, h, 2);
    cases.push({ grid: c.name, id: sdf.id, h, k: 2, count: a.length / 8, offset: off });
    blobs.push(Buffer.from(a.buffer, a.byteOffset, a.byteLength));
    off += a.length;
  }
});
fs.writeFileSync(path.join(dir, 'drip_golden.json'), JSON.stringify({ source: 'dripSamples (src/flip/bake.js:19-40)', cases }));
fs.writeFileSync(path.join(dir, 'drip_golden.bin'), Buffer.concat(blobs));
console.log(cases.map((c) => `${c.grid} h=${c.h}: ${c.count} samples`).join('\n'));

'use strict';
// The facade's buildSDF on the golden meshes (tests/golden/sdf_golden.*): prints
// {name: number of grid values that differ from the reference's bits}.
const fs = require('fs');
const path = require('path');
const { buildSDF } = require(path.join(__dirname, '..', '..', 'splashtank_amd', 'flip'));
const gold = path.join(__dirname, '..', 'golden');
const meta = JSON.parse(fs.readFileSync(path.join(gold, 'sdf_golden.json'), 'utf8'));
const raw = fs.readFileSync(path.join(gold, 'sdf_golden.bin'));
const all = new Uint32Array(raw.buffer, raw.byteOffset, raw.byteLength / 4);
const out = {};
for (const c of meta.cases) {
  const f = [];
  for (let q = 0; q < c.tris.length; q += 3) f.push([c.tris[q], c.tris[q + 1], c.tris[q + 2]]);
  const sd = buildSDF({ v: c.v, f }, c.cell, c.margin);
  const got = new Uint32Array(sd.data.buffer, sd.data.byteOffset, sd.data.length);
  let bad = sd.data.length === c.count ? 0 : -1;
  if (bad === 0) for (let i = 0; i < c.count; i++) if (got[i] !== all[c.offset + i]) bad++;
  out[c.name] = { bad, dims: sd.dims.join('x'), id: sd.id };
}
console.log(JSON.stringify(out));

'use strict';
// The facade's dripSamples on the golden grids: prints {case: number of f32 values that differ}.
const fs = require('fs');
const path = require('path');
const { dripSamples } = require(path.join(__dirname, '..', '..', 'splashtank_amd', 'flip', 'sdf.js'));
const gold = path.join(__dirname, '..', 'golden');
const rd = (n) => { const b = fs.readFileSync(path.join(gold, n)); return new Float32Array(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength)); };
const sm = JSON.parse(fs.readFileSync(path.join(gold, 'sdf_golden.json'), 'utf8')), sb = rd('sdf_golden.bin');
const dm = JSON.parse(fs.readFileSync(path.join(gold, 'drip_golden.json'), 'utf8')), db = rd('drip_golden.bin');
const grids = {};
sm.cases.forEach((c, q) => { grids[c.name] = { id: q + 1, dims: c.dims, origin: c.origin, cell: c.sdfCell, data: sb.subarray(c.offset, c.offset + c.count) }; });
const out = {};
for (const c of dm.cases) {
  const a = dripSamples(grids[c.grid], c.h, c.k), want = db.subarray(c.offset, c.offset + c.count * 8);
  let bad = a.length === want.length ? 0 : -1;
  if (!bad) { const u = new Uint32Array(a.buffer, a.byteOffset, a.length), w = new Uint32Array(want.buffer, want.byteOffset, want.length); for (let i = 0; i < u.length; i++) if (u[i] !== w[i]) bad++; }
  out[c.grid + '@' + c.h] = bad;
}
console.log(JSON.stringify(out));

'use strict';
// The facade's cache writers (splashtank_amd/flip/cache.js) against the bytes the
// reference's own packFrameBuf, packDeep, packSprayJS, packSprayFrame and
// writeCacheParts made from the same inputs (tests/golden/cache_golden.json).
// Prints one JSON line: which parts are equal byte for byte.
const fs = require('fs');
const path = require('path');
const C = require(path.join(__dirname, '..', '..', 'splashtank_amd', 'flip', 'cache.js'));
const G = JSON.parse(fs.readFileSync(path.join(__dirname, '..', 'golden', 'cache_golden.json'), 'utf8'));

const buf = (s) => Buffer.from(s, 'base64');
const f32 = (s) => { const b = buf(s); return new Float32Array(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength)); };
const bytes = (a) => Buffer.from(a.buffer, a.byteOffset, a.byteLength).toString('base64');
const same = (a, s) => (a ? bytes(a) : null) === s;
const sameFrame = (f, g) => f.n === g.n && same(f.pos, g.pos) && same(f.spd, g.spd) && (f.fm ? same(f.fm, g.fm) : g.fm === null) &&
  same(f.objs, g.objs) && (f.sp ? (g.sp && f.sp.n === g.sp.n && same(f.sp.pos, g.sp.pos) && same(f.sp.pk, g.sp.pk)) : g.sp === null) &&
  (f.dp ? (g.dp && f.dp.n === g.dp.n && same(f.dp.pos, g.dp.pos) && same(f.dp.spd, g.dp.spd)) : g.dp === null);

const out = {};
const rb = f32(G.readback);
const frame = C.packFrame(rb, G.n, G.b, G.states, G.vs);
out.frame = sameFrame(frame, G.frame);
const sp = f32(G.spray.p), sv = f32(G.spray.v);
const sd = { n: G.spray.n, p: sp, v: sv };
const full = C.packSprayFrame(sd, G.b, 200000, 3), capped = C.packSprayFrame(sd, G.b, 70, 2.5);
out.sprayFull = full.n === G.spray.full.n && same(full.pos, G.spray.full.pos) && same(full.pk, G.spray.full.pk) && full.alive === G.spray.full.alive;
out.sprayCapped = capped.n === G.spray.capped.n && same(capped.pos, G.spray.capped.pos) && same(capped.pk, G.spray.capped.pk);
const codes = new Uint32Array(G.spray.n);
for (let j = 0; j < G.spray.n; j++) codes[j] = C.packSprayCode(sv[j * 4], sv[j * 4 + 1], sv[j * 4 + 2], Math.round(sv[j * 4 + 3]), sp[j * 4 + 3] / 3);
out.codes = same(codes, G.spray.codes);
const deep = C.packDeepFrame(f32(G.deep.input), G.deep.n, G.b, G.deep.vs);
out.deep = deep.n === G.deep.out.n && same(deep.pos, G.deep.out.pos) && same(deep.spd, G.deep.out.spd);
// the file: the same three frames, our writer
const water = C.packFrame(rb.slice(0, 40 * 4).map((x, k) => (k % 4 === 3 && x >= 64 ? x % 64 : x)), 40, G.b, G.states, 1);
const frames = [Object.assign({}, frame, { dp: deep }), water, Object.assign({}, water, { sp: capped })];
const file = C.writeCache(G.cache.meta, frames);
out.file = file.toString('base64') === G.cache.file;
out.fileBytes = [file.length, buf(G.cache.file).length];
const back = C.readCache(file);
out.readBack = back.frames.length === 3 && back.frames[0].sp.n === frame.sp.n && back.frames[0].dp.n === deep.n &&
  bytes(back.frames[0].dp.pos) === bytes(deep.pos) && bytes(back.frames[2].sp.pk) === bytes(capped.pk) &&
  (back.frames[1].sp ? back.frames[1].sp.n : 0) === (water.sp ? water.sp.n : 0) && !back.frames[1].dp;
console.log(JSON.stringify(out));

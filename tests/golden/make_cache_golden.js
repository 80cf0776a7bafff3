'use strict';
// Writes golden vectors of the reference's bake-cache writers: packFrameBuf and
// packDeep (src/gpu.js:512-560), packSprayJS and packSprayFrame
// (src/flip/bake.js:4-58) and writeCacheParts (src/shared.js:691-709), run on
// seeded synthetic inputs: a readback buffer mixing water (speeds, foam
// levels), dead slots, deep-water points and whitewater codes, positions in and
// out of the box; whitewater pools with every kind, zero and downward
// velocities and dead slots; object poses.  Run once in a container that holds
// the reference:
//   node tests/golden/make_cache_golden.js /root/reference/src tests/golden
// Output: cache_golden.json (inputs and the bytes each function made, base64).
const fs = require('fs');
const vm = require('vm');
const path = require('path');
const { TextEncoder } = require('util');
const [srcDir, outDir] = process.argv.slice(2);
const ctx = vm.createContext({ Math, Float32Array, Float64Array, Int32Array, Uint32Array, Uint16Array, Uint8Array, Array, Map, Object,
  JSON, Number, console, TextEncoder });
for (const f of ['shared.js', 'gpu.js', path.join('flip', 'bake.js')]) vm.runInContext(fs.readFileSync(path.join(srcDir, f), 'utf8'), ctx);
vm.runInContext('this.__api = { packFrameBuf, packDeep, packSprayJS, packSprayFrame, writeCacheParts };', ctx);
const api = ctx.__api;

function rng(seed) {   // splitmix32
  let s = seed >>> 0;
  return () => {
    s = (s + 0x9e3779b9) >>> 0;
    let z = s;
    z = Math.imul(z ^ (z >>> 16), 0x85ebca6b) >>> 0;
    z = Math.imul(z ^ (z >>> 13), 0xc2b2ae35) >>> 0;
    return ((z ^ (z >>> 16)) >>> 0) / 4294967296;
  };
}
const r = rng(12345);
const b = { x0: -0.5, x1: 0.5, y0: 0, y1: 1, z0: -0.5, z1: 0.5 };
const b64 = (a) => Buffer.from(a.buffer, a.byteOffset, a.byteLength).toString('base64');

// readback buffer: 240 slots
const N = 240, buf = new Float32Array(N * 4), U = new Uint32Array(buf.buffer);
for (let i = 0; i < N; i++) {
  const o = i * 4, kind = i < 200 ? Math.floor(r() * 5) : 5;
  buf[o] = -0.6 + 1.2 * r(); buf[o + 1] = -0.05 + 1.1 * r(); buf[o + 2] = -0.6 + 1.2 * r();
  if (kind === 0 || kind === 1) buf[o + 3] = 70 * r() + 64 * Math.floor(16 * r());               // water, speed (some above 63) + foam
  else if (kind === 2) buf[o + 1] = -1e6;                                                          // dead
  else if (kind === 3) buf[o + 3] = 4096 + 70 * r();                                               // deep point
  else if (kind === 4) U[o + 3] = (0x80000000 | Math.floor(r() * 0x7fffffff)) >>> 0;             // whitewater code
  else buf[o + 1] = -1e6;                                                                          // trailing dead slots
}
buf[0] = 0.0; buf[1] = 0.5; buf[2] = 0.0; buf[3] = 2.5;   // a point on the rounding half way
const states = [{ x: 0.1, y: 0.2, z: -0.3, rx: 0.5, yaw: 1.25, rz: -0.75 }, { x: 1, y: 2, z: 3, yaw: 0.5 }];
const frame = api.packFrameBuf(buf, N, b, states, 1.7);

// whitewater pool: 400 slots
const S = 400, sp = new Float32Array(S * 4), sv = new Float32Array(S * 4);
for (let j = 0; j < S; j++) {
  const o = j * 4;
  sp[o] = -0.55 + 1.1 * r(); sp[o + 1] = 1.05 * r(); sp[o + 2] = -0.55 + 1.1 * r();
  sp[o + 3] = r() < 0.3 ? 0 : 4 * r() - 0.2;
  const kind = Math.floor(r() * 4);
  sv[o] = 12 * (r() - 0.5); sv[o + 1] = 12 * (r() - 0.5); sv[o + 2] = 12 * (r() - 0.5);
  if (j % 17 === 0) { sv[o] = 0; sv[o + 1] = 0; sv[o + 2] = 0; }
  if (j % 13 === 0) sv[o + 2] = -Math.abs(sv[o + 2]) - 1;
  sv[o + 3] = kind + 0.4 * r();
}
const sd = { n: S, p: sp, v: sv };
const sprayFull = api.packSprayFrame(sd, b, 200000, 3);
const sprayCapped = api.packSprayFrame(sd, b, 70, 2.5);
const codes = new Uint32Array(S);
for (let j = 0; j < S; j++) codes[j] = api.packSprayJS(sv[j * 4], sv[j * 4 + 1], sv[j * 4 + 2], Math.round(sv[j * 4 + 3]), sp[j * 4 + 3] / 3);

// deep points
const D = 50, dpIn = new Float32Array(D * 4);
for (let i = 0; i < D; i++) {
  dpIn[i * 4] = -0.6 + 1.2 * r(); dpIn[i * 4 + 1] = i % 9 === 0 ? -1e6 : r(); dpIn[i * 4 + 2] = -0.6 + 1.2 * r();
  dpIn[i * 4 + 3] = 4096 + 80 * r() - 5;
}
const deep = api.packDeep(dpIn, D, b, 1.3);

// the file: three frames (everything; water only, no foam; whitewater sample)
const water = api.packFrameBuf(buf.slice(0, 40 * 4).map((x, k) => (k % 4 === 3 && x >= 64 ? x % 64 : x)), 40, b, states, 1);
const f3 = Object.assign({}, water, { sp: sprayCapped });
const cache = { d: 0.0125, fps: 30, ids: [3, 9], b, org: [0, 0, 0], maxN: N, seconds: 2.5, engine: 'flip', scene: 'golden',
  desc: { note: 'synthetic' }, diag: [{ f: 0, alive: 10 }], stableIds: true, scale: 1.7, spdMs: true,
  frames: [Object.assign({}, frame, { dp: deep }), water, f3] };
const parts = api.writeCacheParts(cache);
const file = Buffer.concat(parts.map((p) => Buffer.from(p.buffer, p.byteOffset, p.byteLength)));

const frameOut = (f) => ({ n: f.n, pos: b64(f.pos), spd: b64(f.spd), fm: f.fm ? b64(f.fm) : null, objs: b64(f.objs),
  sp: f.sp ? { n: f.sp.n, pos: b64(f.sp.pos), pk: b64(f.sp.pk), alive: f.sp.alive } : null,
  dp: f.dp ? { n: f.dp.n, pos: b64(f.dp.pos), spd: b64(f.dp.spd) } : null });
const out = {
  source: 'packFrameBuf, packDeep (src/gpu.js:512-560), packSprayJS, packSprayFrame (src/flip/bake.js:4-58), writeCacheParts (src/shared.js:691-709)',
  b, states, readback: b64(buf), n: N, vs: 1.7, frame: frameOut(frame),
  spray: { p: b64(sp), v: b64(sv), n: S, full: frameOut({ n: 0, pos: new Uint16Array(0), spd: new Uint8Array(0), objs: new Float32Array(0), sp: sprayFull }).sp,
    capped: frameOut({ n: 0, pos: new Uint16Array(0), spd: new Uint8Array(0), objs: new Float32Array(0), sp: sprayCapped }).sp, codes: b64(codes) },
  deep: { input: b64(dpIn), n: D, vs: 1.3, out: { n: deep.n, pos: b64(deep.pos), spd: b64(deep.spd) } },
  cache: { meta: Object.assign({}, cache, { frames: undefined }), file: file.toString('base64') },
};
fs.writeFileSync(path.join(outDir, 'cache_golden.json'), JSON.stringify(out));
console.log(`frame: ${frame.n} water, ${frame.sp ? frame.sp.n : 0} whitewater, ${frame.dp ? frame.dp.n : 0} deep; ` +
  `spray ${sprayFull.n}/${sprayCapped.n}; deep ${deep.n}; file ${file.length} bytes`);

'use strict';
// A short bake through runner.bake on the GPU: a dam break, a box collider
// moving along x on keyframes, a water emitter and a dropped ball.  Prints one
// JSON line: per frame the particle count, substeps, and whether every particle
// is finite and inside the box.
const path = require('path');
const M = require(path.join(__dirname, '..', '..', 'splashtank_amd', 'flip'));

// distance grid of a box of half-extent hb sampled on a cell-h grid with a margin
function boxSdf(hb, h, margin) {
  const lo = -hb - margin, n = Math.ceil((2 * (hb + margin)) / h) + 1, data = new Float32Array(n * n * n);
  for (let k = 0; k < n; k++) for (let j = 0; j < n; j++) for (let i = 0; i < n; i++) {
    const p = [lo + i * h, lo + j * h, lo + k * h], q = p.map((v) => Math.abs(v) - hb);
    const out = Math.hypot(Math.max(q[0], 0), Math.max(q[1], 0), Math.max(q[2], 0)), inside = Math.min(Math.max(q[0], q[1], q[2]), 0);
    data[i + n * (j + n * k)] = out + inside;
  }
  return { id: 'box', dims: [n, n, n], origin: [lo, lo, lo], cell: h, data };
}
const d = 1 / 64, h = 2 * d;
const desc = {
  d, maxN: 400000, size: [1, 1, 1], open: 0, ratio: 2, fps: 60, frames: 4, substeps: 2, maxSub: 12, cfl: 1.5, g: 9.81, iters: 20,
  fills: [{ t: 'box', x0: -0.5, x1: -0.2, y0: 0, y1: 0.4, z0: -0.5, z1: 0.5 }],
  objs: [
    { id: 1, kind: 'mesh', sdf: boxSdf(0.08, h, 4 * h), smooth: false, keys: [{ t: 0, x: 0.1, y: 0.15, z: 0 }, { t: 1, x: 0.6, y: 0.15, z: 0 }] },
    { id: 2, kind: 'emitter', on: true, x: 0.2, y: 0.8, z: 0, az: 0, el: -90, radius: 0.05, speed: 1.5 },
  ],
  events: [{ t: 0.03, x: 0.25, y: 0.6, z: 0.2, r: 0.06 }],
};
const frames = [], cache = [];
const box = { x0: -0.5, x1: 0.5, y0: 0, y1: 1, z0: -0.5, z1: 0.5 };
let worstQ = 0;
M.bake(desc, M.initHip(0), (fr, info) => {
  // the frame comes packed as the page's cache holds it; its 16-bit positions decode inside the box
  cache.push(fr);
  for (let i = 0; i < fr.n; i++) {
    if (fr.spd[i] === 255) continue;
    for (let a = 0; a < 3; a++) {
      const lo = [box.x0, box.y0, box.z0][a], ext = [box.x1 - box.x0, box.y1 - box.y0, box.z1 - box.z0][a];
      const x = lo + fr.pos[i * 3 + a] * ext / 65535;
      if (x < lo - 1e-9 || x > lo + ext + 1e-9) worstQ = 1;
    }
  }
  const dg = info.diag;
  frames.push({ f: info.f, n: dg.alive, sub: dg.sub, finite: dg.nonFinite === 0, inside: dg.outsideBox === 0,
    objectSpeed: dg.objectSpeed || 0, maxSpeed: dg.maxSpeed });
}).then((ok) => {
  const file = M.writeCache({ d, fps: desc.fps, ids: desc.objs.map((o) => o.id), b: box, maxN: desc.maxN, engine: 'mi355x-flip' }, cache);
  const back = M.readCache(file);
  const same = back.frames.length === cache.length && back.frames.every((f, k) => f.n === cache[k].n &&
    Buffer.from(f.pos.buffer).equals(Buffer.from(cache[k].pos.buffer)) && Buffer.from(f.spd.buffer).equals(Buffer.from(cache[k].spd.buffer)));
  console.log(JSON.stringify({ ok, frames, cacheBytes: file.length, cacheRoundTrip: same, worstQuant: worstQ }));
})
  .catch((e) => { console.error(e.stack || String(e)); process.exit(1); });

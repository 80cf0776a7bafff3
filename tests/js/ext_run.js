'use strict';
// The extensions through the Node facade: stepN with and without HIP graph
// replay against step() one at a time, the SOR autotuner (state unchanged),
// snapshot/restore, and the adjoint (record, seed, backward, grad).  Reads the
// scene from a JSON + binary pair written by the Python test and writes the
// readbacks and gradients as raw f32 for it to compare with the C ABI.
const fs = require('fs');
const path = require('path');
const { FlipSim, FlipAdjoint, initHip } = require(path.join(__dirname, '..', '..', 'splashtank_amd', 'flip'));

const cfg = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
const raw = fs.readFileSync(cfg.particles);
const arr = new Float32Array(raw.buffer, raw.byteOffset, raw.byteLength / 4);
const n = arr.length / 8;
const dev = initHip(0);
const make = () => {
  const sim = new FlipSim(dev, { d: cfg.d, maxN: n, size: cfg.size, open: cfg.open, ratio: cfg.ratio });
  for (let i = 0; i < n; i++) { const p = i * 8; sim.add(arr[p], arr[p + 1], arr[p + 2], arr[p + 4], arr[p + 5], arr[p + 6]); }
  return sim;
};
const o = Object.assign({ cols: [] }, cfg.params), dt = cfg.params.dt, steps = cfg.steps;
const read = async (sim) => { const dst = new Float32Array(n * 4); await sim.readbackInto(dst); return dst; };
const same = (a, b) => Buffer.from(a.buffer).equals(Buffer.from(b.buffer));

(async () => {
  const out = {};
  const a = make(); for (let s = 0; s < steps; s++) a.step(dt, o);
  const ra = await read(a);
  const b = make(); b.stepN(dt, o, steps, false);
  const c = make(); c.stepN(dt, o, steps, true);
  out.stepN = same(ra, await read(b));
  out.stepNGraph = same(ra, await read(c));
  // the tuner leaves the state as it was: the next step gives the same bits
  const tune = b.tuneSor(dt, o, 2);
  out.tune = { mode: tune.mode, width: tune.width, candidates: tune.candidates.length };
  b.step(dt, o); a.step(dt, o);
  out.afterTune = same(await read(a), await read(b));
  // snapshot and restore: two substeps, back, the same two again
  const snap = a.snapshot();
  a.step(dt, o); a.step(dt, o);
  const r1 = await read(a);
  a.restore(snap); a.dropSnapshot(snap);
  a.step(dt, o); a.step(dt, o);
  out.restore = same(r1, await read(a));
  // the adjoint of `steps` substeps from a fresh start, seeded with dL/dpos = (1, 2, 3) per particle
  const d = make(), adj = new FlipAdjoint(d, steps);
  for (let s = 0; s < steps; s++) adj.record(dt, o);
  const gp = new Float32Array(n * 4), gv = new Float32Array(n * 4);
  for (let i = 0; i < n; i++) { gp[i * 4] = 1; gp[i * 4 + 1] = 2; gp[i * 4 + 2] = 3; }
  adj.seed(gp, gv);
  for (let s = 0; s < steps; s++) adj.backward();
  const g0 = new Float32Array(n * 4), g1 = new Float32Array(n * 4);
  adj.grad(g0, g1);
  fs.writeFileSync(cfg.out, Buffer.concat([Buffer.from(g0.buffer), Buffer.from(g1.buffer)]));
  adj.destroy();
  [a, b, c, d].forEach((s) => s.destroy());
  console.log(JSON.stringify(out));
})().catch((e) => { console.error(e.stack || String(e)); process.exit(1); });

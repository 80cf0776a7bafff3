'use strict';
// Drives FlipBatch over several FlipSim tanks: add() each tank's particles,
// step(dt, [o per tank]) a few times, readbackInto() every tank.  Reads the
// tanks from a JSON file written by the Python test and writes each readback as
// raw f32 for it to compare.
const fs = require('fs');
const path = require('path');
const { FlipSim, FlipBatch, initHip } = require(path.join(__dirname, '..', '..', 'splashtank_amd', 'flip'));

const cfg = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
const dev = initHip(0);
const sims = cfg.tanks.map((t) => {
  const raw = fs.readFileSync(t.particles);
  const arr = new Float32Array(raw.buffer, raw.byteOffset, raw.byteLength / 4);
  const n = arr.length / 8;
  const sim = new FlipSim(dev, { d: t.d, maxN: n, size: t.size, open: t.open, ratio: t.ratio });
  for (let i = 0; i < n; i++) {
    const p = i * 8;
    sim.add(arr[p], arr[p + 1], arr[p + 2], arr[p + 4], arr[p + 5], arr[p + 6]);
  }
  return sim;
});
const batch = new FlipBatch(sims);
const os = cfg.tanks.map((t) => Object.assign({ cols: [] }, t.params));
for (let s = 0; s < cfg.steps; s++) batch.step(cfg.tanks[0].params.dt, os);
batch.sync();
Promise.all(sims.map((sim, k) => {
  const dst = new Float32Array(sim.N * 4);
  return sim.readbackInto(dst).then((m) => { fs.writeFileSync(cfg.tanks[k].out, Buffer.from(dst.buffer)); return m; });
})).then((ns) => {
  console.log(JSON.stringify({ n: ns, stats: sims.map((s) => s.stats.sc) }));
  batch.destroy();
  sims.forEach((s) => s.destroy());
}).catch((e) => { console.error(e.stack || String(e)); process.exit(1); });

'use strict';
// Drives the JS FlipSim facade the way the reference's callers do: add()
// particles, step(dt, o) with collider records, readbackInto().  Reads the
// particles and settings from a JSON + binary pair written by the Python test,
// writes the readback as raw f32 for it to compare.
const fs = require('fs');
const path = require('path');
const { FlipSim, initHip } = require(path.join(__dirname, '..', '..', 'splashtank_amd', 'flip'));

const cfg = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
const raw = fs.readFileSync(cfg.particles);
const arr = new Float32Array(raw.buffer, raw.byteOffset, raw.byteLength / 4);
const n = arr.length / 8;
const sim = new FlipSim(initHip(0), { d: cfg.d, maxN: n, size: cfg.size, open: cfg.open, ratio: cfg.ratio });
for (let i = 0; i < n; i++) {
  const p = i * 8;
  sim.add(arr[p], arr[p + 1], arr[p + 2], arr[p + 4], arr[p + 5], arr[p + 6]);
}
const o = Object.assign({ cols: [] }, cfg.params);
const times = [];
for (let s = 0; s < cfg.steps; s++) {
  const t0 = Date.now();
  sim.step(cfg.params.dt, o);
  times.push(Date.now() - t0);
}
const timed = cfg.timed ? null : null;
sim.stepTimed(cfg.params.dt, o).then((stageMs) => {
  const dst = new Float32Array(n * 4);
  return sim.readbackInto(dst).then((m) => {
    fs.writeFileSync(cfg.out, Buffer.from(dst.buffer));
    const st = sim.stats;
    sim.pump();
    const fr = sim.outFrame, sum = (a) => a.reduce((x, y) => x + y, 0);
    const view = { n: fr.n, pos: sum(fr.pos), spd: sum(fr.spd), fm: sum(fr.fm) };
    console.log(JSON.stringify({ n: m, N: sim.N, nx: sim.nx, tiles: sim.ntiles, scap: sim.scap, stats: st, stages: Object.keys(stageMs),
      fluidCells: st.fluid, mem: sim.memBytes(), norms: sim.solveNorms(), view }));
    sim.destroy();
  });
}).catch((e) => { console.error(e.stack || String(e)); process.exit(1); });
void timed;

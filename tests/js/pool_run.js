'use strict';
// The block pool through the facade: a sim with 6 slots and opt.sort on is
// primed (prime(), sim.js:163-172) and stepped with maybeGrow growing it as
// the reference's policy says; a sim with every tile pooled and opt.sort off
// steps the same particles.  Prints one JSON line: the pool sizes, overflow
// counts and whether the two readbacks are the same bits.
const fs = require('fs');
const path = require('path');
const { FlipSim, initHip } = require(path.join(__dirname, '..', '..', 'splashtank_amd', 'flip'));

const cfg = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
const raw = fs.readFileSync(cfg.particles);
const arr = new Float32Array(raw.buffer, raw.byteOffset, raw.byteLength / 4);
const n = arr.length / 8;
const dev = initHip(0);
const make = (slots) => {
  const sim = new FlipSim(dev, { d: cfg.d, maxN: n, size: cfg.size, open: cfg.open, ratio: cfg.ratio, slots });
  for (let i = 0; i < n; i++) { const p = i * 8; sim.add(arr[p], arr[p + 1], arr[p + 2], arr[p + 4], arr[p + 5], arr[p + 6]); }
  return sim;
};
const o = Object.assign({ cols: [] }, cfg.params), dt = cfg.params.dt;

(async () => {
  const out = {};
  const a = make(6);
  a.opt.sort = true; a.opt.sortEvery = 4;
  out.slotsAtStart = a.scap;
  await a.prime(o);
  out.slotsAfterPrime = a.scap;
  out.primeStats = { sc: a.stats.sc, ovf: a.stats.ovf };
  for (let s = 0; s < cfg.steps; s++) { a.step(dt, o); if (s % 3 === 2) await a.statPromise; }
  await a.statPromise;
  out.slotsAtEnd = a.scap; out.grew = a.grew; out.ovfTotal = a.ovfTotal + a.stats.ovf; out.peak = a.peak;
  const b = make(0);
  for (let s = 0; s < cfg.steps; s++) b.step(dt, o);
  const ra = new Float32Array(n * 4), rb = new Float32Array(n * 4);
  await a.readbackInto(ra); await b.readbackInto(rb);
  out.same = Buffer.from(ra.buffer).equals(Buffer.from(rb.buffer));
  out.fullPool = b.scap;
  a.destroy(); b.destroy();
  console.log(JSON.stringify(out));
})().catch((e) => { console.error(e.stack || String(e)); process.exit(1); });

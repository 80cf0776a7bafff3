'use strict';
// Asynchronous readbacks through the Node facade.  A dam break is stepped many
// times (enqueued, not waited for), then readbackInto starts: the call returns
// at once, the event loop keeps turning while the copy is pending, and the
// Promise resolves with the state at the call even though more substeps were
// enqueued right after it.  readStats, readSpray and readDeep resolve the same
// words and points as the blocking reads.  Prints one JSON line.
const fs = require('fs');
const path = require('path');
const { FlipSim, initHip } = require(path.join(__dirname, '..', '..', 'splashtank_amd', 'flip'));
const addon = require(path.join(__dirname, '..', '..', 'splashtank_amd', 'flip', 'splashtank_hip.node'));

const cfg = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
const raw = fs.readFileSync(cfg.particles);
const arr = new Float32Array(raw.buffer, raw.byteOffset, raw.byteLength / 4);
const n = arr.length / 8;
const dev = initHip(0);
const make = (extra) => {
  const sim = new FlipSim(dev, Object.assign({ d: cfg.d, maxN: n, size: cfg.size, open: cfg.open, ratio: cfg.ratio }, extra || {}));
  for (let i = 0; i < n; i++) { const p = i * 8; sim.add(arr[p], arr[p + 1], arr[p + 2], arr[p + 4], arr[p + 5], arr[p + 6]); }
  return sim;
};
const o = Object.assign({ cols: [] }, cfg.params), dt = cfg.params.dt;
const same = (a, b) => Buffer.from(a.buffer, a.byteOffset, a.byteLength).equals(Buffer.from(b.buffer, b.byteOffset, b.byteLength));

(async () => {
  const out = {};
  // the reference state: `steps` substeps, then a blocking read
  const ref = make();
  for (let s = 0; s < cfg.steps; s++) ref.step(dt, o);
  const want = new Float32Array(n * 4);
  addon.readParticles(ref.handle, want, n);
  // the same substeps, enqueued; the read starts behind them and more substeps follow at once
  const a = make();
  for (let s = 0; s < cfg.steps; s++) a.step(dt, o);
  const dst = new Float32Array(n * 4);
  let settled = false, turns = 0;
  const t0 = process.hrtime.bigint();
  const p = a.readbackInto(dst).then((m) => { settled = true; return m; });
  out.callMs = Number(process.hrtime.bigint() - t0) / 1e6;
  out.pendingAfterCall = !settled;
  for (let s = 0; s < cfg.after; s++) a.step(dt, o);   // enqueued behind the read: must not change its result
  const spin = () => new Promise((r) => setImmediate(r));
  while (!settled) { turns++; await spin(); }
  out.loopTurnsWhilePending = turns;
  out.resolvedN = await p;
  out.totalMs = Number(process.hrtime.bigint() - t0) / 1e6;
  out.stateAtCall = same(dst, want);
  // the control words: readStats against the blocking read, both after the same substeps
  addon.sync(a.handle);
  const st = await a.readStats();
  const ctl = addon.readCtl(a.handle);
  out.stats = st.sc === ctl[4] && st.ovf === ctl[5] && st.fluid === ctl[0] + ctl[1];
  // whitewater slots and deep points: the asynchronous reads against the blocking ones
  const w = make({});
  w.setSprayCap(cfg.sprayCap);
  const ow = Object.assign({}, o, { spr: {} });
  for (let s = 0; s < cfg.sprSteps; s++) w.step(dt, ow);
  const sa = await w.readSpray(), sb = addon.readSpray(w.handle, cfg.sprayCap);
  out.spray = sa === null ? sb === null : (sb !== null && sa.n === sb.n && same(sa.p, sb.p) && same(sa.v, sb.v));
  out.sprayN = sa ? sa.n : 0;
  [ref, a, w].forEach((s) => s.destroy());
  console.log(JSON.stringify(out));
})().catch((e) => { console.error(e.stack || String(e)); process.exit(1); });

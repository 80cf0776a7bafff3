'use strict';
// The facade's scene driving (runner.js, seed.js) against what the reference's
// own keyPose, sampleObjState, objSpeed, SceneRunner, bakeDiag and nbSeed gave on
// the same inputs with the same random draws (tests/golden/runner_golden.json,
// from tests/golden/make_runner_golden.js).  Prints one JSON line: per part,
// whether it is equal, and the first difference when not.
const fs = require('fs');
const path = require('path');
const crypto = require('crypto');
const FLIP = path.join(__dirname, '..', '..', 'splashtank_amd', 'flip');
const R = require(path.join(FLIP, 'runner.js'));
const S = require(path.join(FLIP, 'seed.js'));
const G = JSON.parse(fs.readFileSync(path.join(__dirname, '..', 'golden', 'runner_golden.json'), 'utf8'));

function splitmix(seed) {
  let s = seed >>> 0;
  return () => {
    s = (s + 0x9e3779b9) >>> 0;
    let z = s;
    z = Math.imul(z ^ (z >>> 16), 0x85ebca6b) >>> 0;
    z = Math.imul(z ^ (z >>> 13), 0xc2b2ae35) >>> 0;
    return ((z ^ (z >>> 16)) >>> 0) / 4294967296;
  };
}
// JSON equality of plain data (numbers exactly; undefined fields dropped as JSON drops them)
const canon = (x) => JSON.stringify(x);
const out = { diffs: [] };
const check = (name, a, b) => {
  const ok = canon(a) === canon(b);
  out[name] = (out[name] === undefined ? true : out[name]) && ok;
  if (!ok && out.diffs.length < 5) out.diffs.push({ name, got: canon(a).slice(0, 300), want: canon(b).slice(0, 300) });
};

// keyPose
for (const k of G.keyPose) check('keyPose', R.keyPose(G.keySets[k.set], k.t, k.smooth), k.pose);
// sampleState
for (const s of G.sample) check('sampleState', G.objs.map((o) => R.sampleState(o, s.t)), s.states);
// objectSpeed
for (const c of G.objSpeed) check('objectSpeed', R.objectSpeed(c.a, c.b, c.dt, c.lim), c.v);

// SceneRunner on a recording sim with the same draws
{
  const g = G.runner, sdf = Object.assign({ data: new Float32Array(1) }, g.sdf);
  class Rec {
    constructor(d) { this.d = d; this.h = 2 * d; this.N = 0; this.adds = []; this.steps = []; this.cap = 1e9; }
    add(x, y, z, vx, vy, vz) { this.N++; this.adds.push([x, y, z, vx, vy, vz]); return true; }
    step(dt, o) { this.steps.push({ dt, o: JSON.parse(JSON.stringify(Object.assign({}, o, { cols: o.cols.map((c) => Object.assign({}, c, { sdf: c.sdf.id })) }))) }); }
  }
  const scene = g.scene.map((o) => (o.sdf !== undefined ? Object.assign({}, o, { sdf }) : o));
  const sim = new Rec(g.d), runner = new R.SceneRunner(sim, splitmix(g.rngSeed));
  for (let s = 0; s < g.steps.length; s++) runner.step(g.dt, scene.map((o) => R.sampleState(o, (s + 1) * g.dt)), g.par);
  check('runnerSteps', sim.steps, g.steps);
  check('runnerAdds', sim.adds, g.adds);
  out.runnerCounts = [sim.adds.length, g.adds.length];
}

// bakeDiag
{
  const rb = Float32Array.from(G.bakeDiag.readback, (x) => (x === 'NaN' ? NaN : x));
  check('bakeDiag', R.bakeDiag(rb, G.bakeDiag.n, G.bakeDiag.b, G.bakeDiag.vcap), G.bakeDiag.diag);
  check('bakeDiag', R.bakeDiag(rb, G.bakeDiag.n, G.bakeDiag.b), G.bakeDiag.diagDefault);
}

// nbSeed into a recording sim (the ocean's wave list from the native library, host code)
{
  let addon = null;
  try { addon = require(path.join(FLIP, 'splashtank_hip.node')); } catch (e) { addon = null; }
  for (const c of G.nbSeed) {
    if (c.oc && !addon) { out.nbSeedSkipped = (out.nbSeedSkipped || []).concat(c.name); continue; }
    const sim = { d: c.d, h: c.h, b: c.b, N: 0, cap: c.cap, adds: [], rand: splitmix(4),
      add(x, y, z, vx, vy, vz) { if (this.N >= this.cap) return false; this.N++; this.adds.push([x, y, z, vx, vy, vz]); return true; },
      oceanWaves(oc, g) { return addon.oceanWaves(oc, g, oc.level - this.b.y0); } };
    const res = S.nbSeed(sim, c.fills, { on: true, band: 3 }, c.oc, c.t, c.g, 1);
    const flat = Float64Array.from([].concat(...sim.adds));
    const digest = crypto.createHash('sha256').update(Buffer.from(flat.buffer)).digest('hex');
    check('nbSeed', { res, nbFills: sim.nbFills, nbInit: sim.nbInit, n: sim.N, head: sim.adds.slice(0, 64) },
      { res: c.res, nbFills: c.nbFills, nbInit: c.nbInit, n: c.n, head: c.head });
    check('nbSeedDigest', digest, c.sha256);
  }
}
console.log(JSON.stringify(out));

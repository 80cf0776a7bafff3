'use strict';
// Writes golden vectors of the reference's scene driving: keyPose and
// sampleObjState (src/shared.js:608-633), objSpeed (shared.js:512-523),
// SceneRunner.step with colliders and emitters (shared.js:525-605; collider
// velocities and angular velocities from pose changes, smooth and random
// emitters with spread and turbulence, jumps past colMax) driving a stand-in sim
// that records what it is given, bakeDiag (src/flip/bake.js:60-79) on a synthetic
// readback, and nbSeed (shared.js:805-853) of boxes, a sphere and the ocean into a
// recording sim.  Math.random is replaced by a seeded splitmix32 so the draws
// repeat.  Run once in a container that holds the reference:
//   node tests/golden/make_runner_golden.js /root/reference/src tests/golden
// Output: runner_golden.json.
const fs = require('fs');
const vm = require('vm');
const path = require('path');
const [srcDir, outDir] = process.argv.slice(2);

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
const SeededMath = Object.create(Math);
SeededMath.random = splitmix(1);
const ctx = vm.createContext({ Math: SeededMath, Float32Array, Float64Array, Int32Array, Uint32Array, Uint16Array, Uint8Array, Array, Map,
  Object, JSON, Number, console });
for (const f of ['shared.js', path.join('flip', 'bake.js')]) vm.runInContext(fs.readFileSync(path.join(srcDir, f), 'utf8'), ctx);
vm.runInContext('this.__api = { keyPose, sampleObjState, objSpeed, SceneRunner, bakeDiag, nbSeed, oceanWaves };', ctx);
const api = ctx.__api;
const out = { source: 'keyPose, sampleObjState, objSpeed, SceneRunner, nbSeed (src/shared.js), bakeDiag (src/flip/bake.js)', rng: 'splitmix32' };

// keyPose
const keySets = [
  [{ t: 0, x: 0, y: 1, z: -1, rx: 0, yaw: 0, rz: 0.1 }, { t: 0.5, x: 1, y: 1.5, z: 0, rx: 0.3, yaw: 1, rz: 0 },
    { t: 1.25, x: 3, y: 0.5, z: 0.25, rx: -0.2, yaw: 2, rz: 0.4 }, { t: 2, x: 3.5, y: 0.75, z: 1, rx: 0, yaw: 2.5, rz: 0 }],
  [{ t: 0.1, x: 0.2, on: true, speed: 1 }, { t: 0.9, x: -0.4, on: false, speed: 3 }],
  [{ t: 0, x: 0, yaw: 0 }, { t: 1, x: 1 }, { t: 2, x: 3, yaw: 1 }],
];
const times = [-0.5, 0, 0.05, 0.1, 0.3, 0.5, 0.77, 1, 1.25, 1.6, 1.99, 2, 3];
out.keyPose = [];
keySets.forEach((keys, si) => {
  for (const smooth of [true, false]) for (const t of times) out.keyPose.push({ set: si, smooth, t, pose: api.keyPose(keys, t, smooth) });
});
out.keySets = keySets;

// sampleObjState
const objs = [
  { id: 1, kind: 'mesh', keys: keySets[0], smooth: true, x: 9, y: 9, z: 9 },
  { id: 2, kind: 'mesh', keys: keySets[0], smooth: false },
  { id: 3, kind: 'emitter', keys: keySets[1], on: true, speed: 0.5 },
  { id: 4, kind: 'mesh', x: 0.1, y: 0.2, z: 0.3, yaw: 0.4, spin: 1.5 },
  { id: 5, kind: 'mesh', x: 0, y: 0, z: 0, rec: { start: 6, arr: [0, 0, 0, 0, 0.1, 0.2, 0.3, 0.4, 0.3, 0.1, -0.2, 0.5, 0.5, 0.5, 0.5, 1] } },
];
out.objs = objs;
out.sample = [];
for (const t of [0, 0.04, 0.1, 0.25, 0.5, 1.3, 5]) out.sample.push({ t, states: objs.map((o) => api.sampleObjState(o, t)) });

// objSpeed
const sdfA = { id: 7, dims: [10, 12, 14], origin: [0, 0, 0], cell: 0.025, data: new Float32Array(1) };
const a0 = [{ id: 1, kind: 'mesh', x: 0, y: 0, z: 0, rx: 0, yaw: 0, rz: 0, sdf: sdfA }, { id: 2, kind: 'mesh', x: 1, y: 1, z: 1 },
  { id: 3, kind: 'emitter', on: true, speed: 0.7 }, { id: 4, kind: 'mesh', x: 0, y: 0, z: 0 }];
const a1 = [{ id: 1, kind: 'mesh', x: 0.01, y: -0.02, z: 0.005, rx: 0.01, yaw: -0.02, rz: 0.03, sdf: sdfA }, { id: 2, kind: 'mesh', x: 1.3, y: 1, z: 1 },
  { id: 3, kind: 'emitter', on: true, speed: 0.7 }, { id: 9, kind: 'mesh', x: 5, y: 0, z: 0 }];
out.objSpeed = [[a0, a1, 0.01, 0], [a0, a1, 0.01, 20], [a0, a1, 0.05, 4], [null, a1, 0.01, 20]].map(([a, b, dt, lim]) =>
  ({ a, b, dt, lim, v: api.objSpeed(a, b, dt, lim) }));

// SceneRunner.step on a recording sim
class Rec {
  constructor(d) { this.d = d; this.h = 2 * d; this.N = 0; this.adds = []; this.steps = []; this.cap = 1e9; }
  add(x, y, z, vx, vy, vz) { this.N++; this.adds.push([x, y, z, vx, vy, vz]); return true; }
  step(dt, o) { this.steps.push({ dt, o: JSON.parse(JSON.stringify(Object.assign({}, o, { cols: o.cols.map((c) => Object.assign({}, c, { sdf: c.sdf.id })) }))) }); }
}
SeededMath.random = splitmix(2);
const sdfB = { id: 11, dims: [8, 8, 8], origin: [-0.1, -0.1, -0.1], cell: 0.025, data: new Float32Array(1) };
const scene = [
  { id: 1, kind: 'mesh', sdf: sdfB, smooth: true, keys: [{ t: 0, x: -0.3, y: 0.2, z: 0, rx: 0, yaw: 0, rz: 0 }, { t: 0.05, x: -0.25, y: 0.22, z: 0.01, rx: 0.05, yaw: 0.1, rz: -0.02 },
    { t: 0.1, x: -0.15, y: 0.2, z: 0.0, rx: 0.1, yaw: 0.3, rz: 0 }] },
  { id: 2, kind: 'mesh', sdf: sdfB, x: 0, y: 0.5, z: 0, water: true },
  { id: 3, kind: 'mesh', sdf: sdfB, keys: [{ t: 0, x: 0 }, { t: 0.01, x: 0 }, { t: 0.02, x: 3 }], smooth: false },
  { id: 4, kind: 'emitter', on: true, x: 0.1, y: 0.6, z: -0.1, az: 30, el: -60, radius: 0.04, speed: 2, spread: 0.2, turb: 0.7 },
  { id: 5, kind: 'emitter', on: true, mode: 'random', x: -0.2, y: 0.7, z: 0.2, az: 120, el: 89, radius: 0.05, speed: 1.5, spread: 0.1, turb: 0.5, density: 0.8 },
  { id: 6, kind: 'emitter', on: true, x: 0.3, y: 0.5, z: 0.3, az: 0, el: 0, radius: 0.03, speed: 1, spread: 0 },
];
const sim = new Rec(1 / 96), runner = new api.SceneRunner(sim), dt = 1 / 240;
for (let s = 0; s < 8; s++) runner.step(dt, scene.map((o) => api.sampleObjState(o, (s + 1) * dt)), { g: 9.81, colMax: 15, scale: 1.3, iters: 20 });
out.runner = { scene: scene.map((o) => Object.assign({}, o, o.sdf ? { sdf: o.sdf.id } : {})), sdf: { id: sdfB.id, dims: sdfB.dims, origin: sdfB.origin, cell: sdfB.cell },
  d: sim.d, dt, par: { g: 9.81, colMax: 15, scale: 1.3, iters: 20 }, steps: sim.steps, adds: sim.adds, rngSeed: 2 };

// bakeDiag
SeededMath.random = splitmix(3);
const rb = new Float32Array(300 * 4);
for (let i = 0; i < 300; i++) {
  rb[i * 4] = -0.6 + 1.2 * Math.random(); rb[i * 4 + 1] = -0.1 + 1.2 * Math.random(); rb[i * 4 + 2] = -0.6 + 1.2 * Math.random();
  rb[i * 4 + 3] = 64 * Math.floor(16 * Math.random()) + 70 * Math.random();
  if (i % 23 === 0) rb[i * 4 + 1] = -1e6;
  if (i % 41 === 0) rb[i * 4] = NaN;
}
const bbox = { x0: -0.5, x1: 0.5, y0: 0, y1: 1, z0: -0.5, z1: 0.5 };
out.bakeDiag = { readback: Array.from(rb, (x) => (Number.isNaN(x) ? 'NaN' : x)), n: 300, b: bbox, vcap: 14, diag: api.bakeDiag(rb, 300, bbox, 14),
  diagDefault: api.bakeDiag(rb, 300, bbox) };

// nbSeed into a recording sim: a box on the floor and walls, a box in the air, a sphere, and the ocean
class SeedRec {
  constructor(d, h, b, cap) { this.d = d; this.h = h; this.b = b; this.N = 0; this.cap = cap; this.adds = []; }
  add(x, y, z, vx, vy, vz) { if (this.N >= this.cap) return false; this.N++; this.adds.push([x, y, z, vx, vy, vz]); return true; }
}
const seeds = [];
for (const [name, fills, oc, cap] of [
  ['boxes_sphere', [{ t: 'box', x0: -0.5, x1: 0.5, y0: 0, y1: 0.3, z0: -0.5, z1: 0.5 }, { t: 'box', x0: -0.2, x1: 0.1, y0: 0.5, y1: 0.7, z0: -0.1, z1: 0.2 },
    { t: 'sphere', x: 0.2, y: 0.6, z: -0.1, r: 0.16, vy: -1.5 }], null, 1e9],
  ['ocean', [{ t: 'sphere', x: 0.1, y: 0.7, z: 0, r: 0.12 }], { on: true, height: 0.04, length: 0.45, dir: 30, spread: 40, seed: 7, level: 0.35, layer: 0.12 }, 1e9],
  ['full', [{ t: 'box', x0: -0.5, x1: 0.5, y0: 0, y1: 0.3, z0: -0.5, z1: 0.5 }], null, 1000],
]) {
  SeededMath.random = splitmix(4);
  const d = 1 / 24, h = 2 * d, b = { x0: -0.5, x1: 0.5, y0: 0, y1: 1, z0: -0.5, z1: 0.5 };
  const s = new SeedRec(d, h, b, cap);
  const res = api.nbSeed(s, fills, { on: true, band: 3 }, oc, 0.4, 9.81, 1);
  // the particles as a digest: sha256 of their f64 values (x y z vx vy vz per particle), and the first 64 in full
  const flat = Float64Array.from([].concat(...s.adds));
  const digest = require('crypto').createHash('sha256').update(Buffer.from(flat.buffer)).digest('hex');
  seeds.push({ name, d, h, b, cap, fills, oc, t: 0.4, g: 9.81, res, nbFills: s.nbFills, nbInit: s.nbInit, n: s.N, sha256: digest,
    head: s.adds.slice(0, 64) });
}
out.nbSeed = seeds;
fs.writeFileSync(path.join(outDir, 'runner_golden.json'), JSON.stringify(out));
console.log(`keyPose ${out.keyPose.length}, runner steps ${sim.steps.length}, adds ${sim.adds.length}, nbSeed ${seeds.map((q) => q.n).join('/')}`);

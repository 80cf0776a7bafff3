'use strict';
// Checks of the scene driving (runner.js) against a stand-in sim that records
// what it is given; no GPU.  Prints one JSON line of results.
const path = require('path');
const R = require(path.join(__dirname, '..', '..', 'splashtank_amd', 'flip', 'runner.js'));

const out = {};
class Rec {
  constructor(d) { this.d = d; this.h = 2 * d; this.N = 0; this.added = []; this.steps = []; }
  add(x, y, z, vx, vy, vz) { this.N++; this.added.push([x, y, z, vx, vy, vz]); return true; }
  step(dt, o) { this.steps.push({ dt, cols: o.cols.map((c) => Object.assign({}, c)) }); return o.cols.length; }
}
const sdf = { id: 'box', dims: [8, 8, 8], origin: [-0.1, -0.1, -0.1], cell: 0.025, data: new Float32Array(512) };

// a box moving along x at 0.5 m/s and turning about y at 2 rad/s
{
  const sim = new Rec(1 / 128), run = new R.SceneRunner(sim), dt = 1 / 240;
  const obj = { id: 1, kind: 'mesh', sdf, keys: [{ t: 0, x: -0.3, y: 0.2, z: 0, yaw: 0 }, { t: 10, x: 4.7, y: 0.2, z: 0, yaw: 20 }], smooth: false };
  for (let s = 0; s < 5; s++) run.step(dt, [R.sampleState(obj, (s + 1) * dt)], { g: 9.81 });
  const c0 = sim.steps[0].cols[0], c4 = sim.steps[4].cols[0];
  out.firstVel = [c0.vx, c0.wy];
  out.vel = [c4.vx, c4.vy, c4.vz];
  out.omega = [c4.wx, c4.wy, c4.wz];
  out.omegaExpect = Math.sin(2 * dt) / dt;
  // a jump faster than colMax gives no velocity
  const jump = { id: 2, kind: 'mesh', sdf, x: 0, y: 0.2, z: 0 };
  const sim2 = new Rec(1 / 128), run2 = new R.SceneRunner(sim2);
  run2.step(dt, [jump], {});
  run2.step(dt, [Object.assign({}, jump, { x: 0.5 })], {});
  out.jumpVel = sim2.steps[1].cols[0].vx;
}
// a smooth emitter: one lattice disc per spacing the stream advances
{
  const d = 1 / 64, sim = new Rec(d), run = new R.SceneRunner(sim), dt = 1 / 240;
  const em = { id: 3, kind: 'emitter', on: true, x: 0, y: 0.5, z: 0, az: 0, el: -90, radius: 0.05, speed: 2 };
  const steps = 48;
  for (let s = 0; s < steps; s++) run.step(dt, [em], {});
  let disc = 0; const Rr = Math.ceil(em.radius / d) + 1;
  for (let i = -Rr; i <= Rr; i++) for (let j = -Rr; j <= Rr; j++) { const a = (i + 0.5) * d, c = (j + 0.5) * d; if (a * a + c * c <= em.radius * em.radius) disc++; }
  out.emitted = sim.N;
  out.emittedExpect = Math.floor(steps * dt * em.speed / d + 1e-9) * disc;
  out.emitVel = sim.added[0].slice(3);
}
// keyframed poses
{
  const keys = [{ t: 0, x: 0 }, { t: 1, x: 1 }, { t: 2, x: 3 }, { t: 3, x: 3 }];
  out.keyHeld = [R.keyPose(keys, -1, true).x, R.keyPose(keys, 5, true).x];
  out.keyLinear = R.keyPose(keys, 1.5, false).x;
  out.keySpline = [R.keyPose(keys, 1, true).x, R.keyPose(keys, 2, true).x, R.keyPose(keys, 1.5, true).x];
  out.objSpeed = R.objectSpeed([{ id: 1, kind: 'mesh', x: 0, y: 0, z: 0 }], [{ id: 1, kind: 'mesh', x: 0.01, y: 0, z: 0 }], 0.01, 20);
}
console.log(JSON.stringify(out));

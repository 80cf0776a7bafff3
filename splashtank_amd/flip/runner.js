'use strict';
// Scene driving on top of FlipSim: the caller side of the substep.
//
// SceneRunner.step(dt, states, par) turns the scene's objects at one instant
// into a FlipSim.step: mesh objects with a distance grid become collider
// records whose velocity and angular velocity come from the change of pose
// since the previous step (as the reference's SceneRunner does,
// src/shared.js:525-550), and emitters release water (shared.js:551-604).
// keyPose / sampleState give an object's pose at time t from its keyframes
// (shared.js:608-633), and bake() is the adaptive-substep bake loop of
// src/flip/bake.js:81-132: each frame is cut into enough substeps that the
// fastest particle or object moves at most `cfl` cells per substep.
//
// Differences, each deliberate: emitter jitter comes from a seeded generator
// (so a bake repeats exactly) and emitters have no turbulence field; frames are
// handed over as positions (the SPLC packing is a separate format).

// R = Rz(rz) Ry(ry) Rx(rx), row major
function eulerRows(rx, ry, rz) {
  const cx = Math.cos(rx), sx = Math.sin(rx), cy = Math.cos(ry), sy = Math.sin(ry), cz = Math.cos(rz), sz = Math.sin(rz);
  return [cz * cy, cz * sy * sx - sz * cx, cz * sy * cx + sz * sx,
    sz * cy, sz * sy * sx + cz * cx, sz * sy * cx - cz * sx,
    -sy, cy * sx, cy * cx];
}

// pose at time t from keyframes sorted by t: held before the first and after
// the last key; in between linear, or a Catmull-Rom spline through the keys
function keyPose(keys, t, smooth) {
  const n = keys.length;
  if (t <= keys[0].t) return keys[0];
  if (t >= keys[n - 1].t) return keys[n - 1];
  let i = 0;
  while (i < n - 2 && t >= keys[i + 1].t) i++;
  const a = keys[i], b = keys[i + 1], u = (t - a.t) / Math.max(b.t - a.t, 1e-6);
  const p0 = i > 0 ? keys[i - 1] : a, p3 = i + 2 < n ? keys[i + 2] : b;
  const at = (k) => {
    const A = a[k] || 0, B = b[k] || 0;
    if (!smooth) return A + (B - A) * u;
    const P = p0[k] || 0, Q = p3[k] || 0;
    // Catmull-Rom: tangents (B - P) / 2 at A and (Q - A) / 2 at B
    return A + 0.5 * u * ((B - P) + u * ((2 * P - 5 * A + 4 * B - Q) + u * (3 * (A - B) + Q - P)));
  };
  const sp = a.speed === undefined ? undefined : a.speed + ((b.speed === undefined ? a.speed : b.speed) - a.speed) * u;
  return { t, x: at('x'), y: at('y'), z: at('z'), rx: at('rx'), yaw: at('yaw'), rz: at('rz'), on: a.on, speed: sp };
}

function sampleState(o, t) {
  const s = Object.assign({}, o);
  if (o.keys && o.keys.length) {
    const k = keyPose(o.keys, t, o.smooth !== false);
    s.x = k.x; s.y = k.y; s.z = k.z; s.rx = k.rx || 0; s.yaw = k.yaw || 0; s.rz = k.rz || 0;
    if (o.kind === 'emitter') { if (k.on !== undefined) s.on = k.on; if (k.speed !== undefined) s.speed = k.speed; }
  }
  return s;
}

// fastest object motion between two samples of the same objects (translation
// plus rotation times the object's half-diagonal), jumps above `lim` ignored
function objectSpeed(a, b, dt, lim) {
  let v = 0;
  for (let i = 0; i < b.length; i++) {
    const o = b[i], p = a && a[i];
    if (o.kind === 'emitter') { if (o.on) v = Math.max(v, o.speed || 0); continue; }
    if (!p || p.id !== o.id) continue;
    const lin = Math.hypot(o.x - p.x, o.y - p.y, o.z - p.z);
    const rot = Math.abs((o.rx || 0) - (p.rx || 0)) + Math.abs((o.yaw || 0) - (p.yaw || 0)) + Math.abs((o.rz || 0) - (p.rz || 0));
    const half = o.sdf ? 0.5 * o.sdf.cell * Math.hypot(o.sdf.dims[0], o.sdf.dims[1], o.sdf.dims[2]) : 0.5;
    const ov = (lin + rot * half) / dt;
    if (!lim || ov <= lim) v = Math.max(v, ov);
  }
  return v;
}

// splitmix32: the emitters' jitter
function rng32(seed) {
  let s = seed >>> 0;
  return () => {
    s = (s + 0x9e3779b9) >>> 0;
    let z = s;
    z = Math.imul(z ^ (z >>> 16), 0x85ebca6b) >>> 0;
    z = Math.imul(z ^ (z >>> 13), 0xc2b2ae35) >>> 0;
    return ((z ^ (z >>> 16)) >>> 0) / 4294967296;
  };
}

class SceneRunner {
  constructor(sim, seed) {
    this.sim = sim;
    this.prev = new Map();   // object id -> last pose (x, y, z, R)
    this.em = new Map();     // emitter id -> release state
    this.cols = [];
    this.rand = rng32(seed === undefined ? 1 : seed);
  }
  // collider record of a mesh object, velocities from the pose change over dt
  collider(o, dt, lim2) {
    const R = eulerRows(o.rx || 0, o.yaw || 0, o.rz || 0), p = this.prev.get(o.id);
    this.prev.set(o.id, { x: o.x, y: o.y, z: o.z, R });
    const c = { sdf: o.sdf, x: o.x, y: o.y, z: o.z, R, vx: 0, vy: 0, vz: 0, wx: 0, wy: 0, wz: 0 };
    if (!p) return c;
    const vx = (o.x - p.x) / dt, vy = (o.y - p.y) / dt, vz = (o.z - p.z) / dt;
    if (vx * vx + vy * vy + vz * vz < lim2) { c.vx = vx; c.vy = vy; c.vz = vz; }
    // dR = R Rprev^T ~ I + [w]x dt: the angular velocity from its skew part
    const P = p.R, d = new Array(9);
    for (let r = 0; r < 3; r++)
      for (let q = 0; q < 3; q++) d[r * 3 + q] = R[r * 3] * P[q * 3] + R[r * 3 + 1] * P[q * 3 + 1] + R[r * 3 + 2] * P[q * 3 + 2];
    const wx = (d[7] - d[5]) / (2 * dt), wy = (d[2] - d[6]) / (2 * dt), wz = (d[3] - d[1]) / (2 * dt);
    if (wx * wx + wy * wy + wz * wz < lim2) { c.wx = wx; c.wy = wy; c.wz = wz; }
    return c;
  }
  step(dt, states, par) {
    const lim2 = Math.pow(par.colMax || 20, 2);
    this.scale = par.scale || 1;
    const cols = [];
    for (const o of states) {
      if (o.kind === 'mesh') { if (o.sdf && !o.water) cols.push(this.collider(o, dt, lim2)); }
      else if (o.kind === 'emitter' && o.on) this.emit(o, dt);
    }
    this.cols = cols;
    return this.sim.step(dt, Object.assign({}, par, { cols }));
  }
  // an emitter: a disc of radius o.radius at (x, y, z) facing azimuth o.az,
  // elevation o.el (degrees), releasing water at o.speed
  emit(o, dt) {
    const sim = this.sim, d = sim.d, rnd = this.rand;
    let st = this.em.get(o.id);
    if (!st) { st = { phase: 0, acc: 0, key: '', pts: null }; this.em.set(o.id, st); }
    const az = o.az * Math.PI / 180, el = o.el * Math.PI / 180;
    const dx = Math.cos(el) * Math.cos(az), dy = Math.sin(el), dz = Math.cos(el) * Math.sin(az);
    const ref = Math.abs(dy) > 0.95 ? [1, 0, 0] : [0, 1, 0];
    let ux = dy * ref[2] - dz * ref[1], uy = dz * ref[0] - dx * ref[2], uz = dx * ref[1] - dy * ref[0];
    const ul = Math.hypot(ux, uy, uz) || 1; ux /= ul; uy /= ul; uz /= ul;
    const wx = dy * uz - dz * uy, wy = dz * ux - dx * uz, wz = dx * uy - dy * ux;
    const ks = 1 / this.scale, sp = (o.spread || 0) * ks, v = o.speed * ks, veff = Math.max(v, 0.3 * ks);
    const jit = () => (sp ? (rnd() - 0.5) * 2 * sp : 0);
    const put = (a, c, adv) => sim.add(o.x + ux * a + wx * c + dx * adv, o.y + uy * a + wy * c + dy * adv, o.z + uz * a + wz * c + dz * adv,
      dx * v + jit(), dy * v + jit(), dz * v + jit());
    if (o.mode === 'random') {   // a Poisson-like stream: the disc's volume flux in particles
      st.acc += Math.PI * o.radius * o.radius * veff / (d * d * d) * (o.density || 1) * dt;
      const n = Math.floor(st.acc);
      st.acc -= n;
      for (let k = 0; k < n; k++) {
        const a = rnd() * Math.PI * 2, rr = o.radius * Math.sqrt(rnd());
        if (!put(Math.cos(a) * rr, Math.sin(a) * rr, rnd() * veff * dt)) break;
      }
      return;
    }
    // smooth: a square-lattice disc each time the stream has advanced one spacing
    const key = o.radius + '|' + d;
    if (st.key !== key) {
      const pts = [], R = Math.ceil(o.radius / d) + 1;
      for (let i = -R; i <= R; i++)
        for (let j = -R; j <= R; j++) {
          const a = (i + 0.5) * d, c = (j + 0.5) * d;
          if (a * a + c * c <= o.radius * o.radius) pts.push(a, c);
        }
      if (!pts.length) pts.push(0, 0);
      st.pts = pts; st.key = key;
    }
    st.phase += veff * dt;
    for (let guard = 0; st.phase >= d && guard < 8; guard++) {
      st.phase -= d;
      for (let k = 0; k < st.pts.length; k += 2) if (!put(st.pts[k], st.pts[k + 1], st.phase)) return;
    }
  }
}

// The bake loop: desc = { d, maxN, size, open, ratio, fills: [{t:'box', x0..z1} |
// {t:'sphere', x, y, z, r, vy}], fps, frames, substeps (least per frame),
// maxSub, cfl, g, iters, flip, comp, objs (scene objects), events ([{t, x, y, z,
// r}] balls of water dropped at time t), colMax }.  onFrame(frame, info) gets
// every frame's positions (id order, xyz + speed and foam in w) and the substep
// count used.  Resolves true when all frames are done.
async function bake(desc, device, onFrame, isCancelled) {
  const { FlipSim } = require('./sim.js');
  const sim = new FlipSim(device, { d: desc.d, maxN: desc.maxN, size: desc.size, open: desc.open | 0, ratio: desc.ratio });
  await sim.whenReady;
  sim.recycle = false;
  const runner = new SceneRunner(sim, desc.seed);
  for (const f of desc.fills || []) {
    if (f.t === 'box') sim.fillBox(f.x0, f.x1, f.y0, f.y1, f.z0, f.z1);
    else sim.fillSphere(f.x, f.y, f.z, f.r, 0, f.vy || 0, 0);
  }
  const objs = desc.objs || [];
  const all = (t) => objs.map((o) => sampleState(o, t));
  const dtF = 1 / desc.fps, F = desc.frames, minSub = desc.substeps || 1, cfl = desc.cfl || 1.5, maxSub = Math.max(minSub, desc.maxSub || 16);
  const par = { g: desc.g === undefined ? 9.81 : desc.g, iters: desc.iters || 30, flip: desc.flip, comp: desc.comp, colMax: desc.colMax };
  const ev = (desc.events || []).slice().sort((a, b) => a.t - b.t);
  let ei = 0, vmax = 2;
  for (const o of all(0)) if (o.kind === 'emitter') vmax = Math.max(vmax, o.speed || 0);
  if (sim.N > 0) await sim.prime(par);
  let dst = new Float32Array(Math.max(sim.N, 1) * 4);
  let n = await sim.readbackInto(dst);
  onFrame({ n, pos: dst.slice(0, n * 4) }, { f: 0, F, n, sub: 0 });
  let ok = true;
  for (let f = 1; f <= F; f++) {
    const objV = objectSpeed(all((f - 1) * dtF), all(f * dtF), dtF, desc.colMax);
    const sub = Math.max(minSub, Math.min(maxSub, Math.ceil(dtF * Math.max(vmax, objV) / (cfl * sim.h))));
    const dt = dtF / sub;
    for (let s = 0; s < sub; s++) {
      const t = (f - 1) * dtF + (s + 1) * dt;
      while (ei < ev.length && ev[ei].t <= t) { const e = ev[ei++]; sim.fillSphere(e.x, e.y, e.z, e.r, 0, -1, 0); }
      runner.step(dt, all(t), par);
    }
    if (dst.length < sim.N * 4) dst = new Float32Array(sim.N * 4);
    n = await sim.readbackInto(dst);
    // the next frame's substeps follow this frame's fastest particle (w holds
    // the speed below 64) and the emitters
    let mv = 0;
    for (let i = 0; i < n; i++) { const w = dst[i * 4 + 3] % 64; if (w > mv) mv = w; }
    const st = sim.stats;
    if (st && st.vmax > mv) mv = st.vmax;
    for (const o of all(f * dtF)) if (o.kind === 'emitter' && o.on) mv = Math.max(mv, o.speed || 0);
    vmax = Math.max(1, mv);
    onFrame({ n, pos: dst.slice(0, n * 4) }, { f, F, n, sub, objectSpeed: objV, fastest: st ? st.vmax : 0 });
    if (isCancelled && isCancelled()) { ok = false; break; }
  }
  sim.destroy();
  return ok;
}

module.exports = { SceneRunner, bake, keyPose, sampleState, objectSpeed, eulerRows };

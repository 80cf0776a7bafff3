'use strict';
// Scene driving on top of FlipSim: the caller side of the substep.
//
// SceneRunner.step(dt, states, par) turns the scene's objects at one instant
// into a FlipSim.step: mesh objects with a distance grid become collider
// records whose velocity and angular velocity come from the change of pose
// since the previous step (the reference's SceneRunner, src/shared.js:525-550),
// and emitters release water, optionally with turbulent eddies (shared.js:551-604).
// keyPose / sampleState give an object's pose at time t from its keyframes, a
// recorded path or a constant spin (shared.js:608-633), objectSpeed the fastest
// object motion between two instants (objSpeed, shared.js:512-523), bakeDiag the
// per-frame report numbers (src/flip/bake.js:60-79), and bake() is the
// adaptive-substep bake loop (runBakeFlipGpu, src/flip/bake.js:81-132): each frame
// is cut into enough substeps that the fastest particle or object moves at most
// `cfl` cells per substep.
//
// Random numbers (emitter jitter and turbulence) come from the runner's `rand`,
// Math.random unless the caller passes a seeded generator; they are drawn in
// the reference's order, so one generator gives the reference's particles.

const { nbSeed, fillVolume } = require('./seed.js');
const { packFrame, packSprayFrame, packDeepFrame } = require('./cache.js');

// R = Rz(rz) Ry(ry) Rx(rx), row major
function eulerRows(rx, ry, rz) {
  const cx = Math.cos(rx), sx = Math.sin(rx), cy = Math.cos(ry), sy = Math.sin(ry), cz = Math.cos(rz), sz = Math.sin(rz);
  return [cz * cy, cz * sy * sx - sz * cx, cz * sy * cx + sz * sx,
    sz * cy, sz * sy * sx + cz * cx, sz * sy * cx - cz * sx,
    -sy, cy * sx, cy * cx];
}

// One coordinate of a pose between keys a and b at u in [0, 1]: linear, or the
// Catmull-Rom cubic through the neighbouring keys p0 and p3, as the polynomial
// 0.5 (2A + (B - P0) u + (2 P0 - 5A + 4B - P3) u^2 + (3A - P0 - 3B + P3) u^3)
// evaluated term by term in that order.
function keyCoord(p0, a, b, p3, k, u, smooth) {
  const A = a[k] || 0, B = b[k] || 0;
  if (!smooth) return A + (B - A) * u;
  const P = p0[k] || 0, Q = p3[k] || 0, u2 = u * u, u3 = u2 * u;
  const c1 = -P + B, c2 = 2 * P - 5 * A + 4 * B - Q, c3 = -P + 3 * A - 3 * B + Q;
  return 0.5 * ((2 * A) + c1 * u + c2 * u2 + c3 * u3);
}
// pose at time t from keyframes sorted by t: held before the first and after the
// last key; emitters carry their on/off switch and speed along
function keyPose(keys, t, smooth) {
  const last = keys.length - 1;
  if (t <= keys[0].t) return keys[0];
  if (t >= keys[last].t) return keys[last];
  let i = 0;
  while (i < last - 1 && t >= keys[i + 1].t) i++;
  const a = keys[i], b = keys[i + 1], u = (t - a.t) / Math.max(b.t - a.t, 1e-6);
  const p0 = keys[i - 1] || a, p3 = keys[i + 2] || b;
  const out = { t };
  for (const k of ['x', 'y', 'z', 'rx', 'yaw', 'rz']) out[k] = keyCoord(p0, a, b, p3, k, u, smooth);
  out.on = a.on;
  out.speed = a.speed === undefined ? undefined : a.speed + (b.speed - a.speed) * u;
  return out;
}

// an object's state at time t: keyframes, else a path recorded at 60 samples per
// second (rec = {start, arr: x y z yaw per sample}), else a constant spin about y
function sampleState(o, t) {
  const s = Object.assign({}, o);
  if (o.keys && o.keys.length) {
    const k = keyPose(o.keys, t, o.smooth !== false);
    s.x = k.x; s.y = k.y; s.z = k.z; s.yaw = k.yaw || 0; s.rx = k.rx || 0; s.rz = k.rz || 0;
    if (o.kind === 'emitter') {
      if (k.on !== undefined) s.on = k.on;
      if (k.speed !== undefined) s.speed = k.speed;
    }
    return s;
  }
  if (o.rec && o.rec.arr.length >= 4) {
    const a = o.rec.arr, n = a.length / 4;
    const at = Math.min(Math.max(t * 60 - o.rec.start, 0), n - 1);
    const i0 = Math.floor(at), i1 = Math.min(i0 + 1, n - 1), w = at - i0;
    const lerp = (q) => a[i0 * 4 + q] + (a[i1 * 4 + q] - a[i0 * 4 + q]) * w;
    s.x = lerp(0); s.y = lerp(1); s.z = lerp(2); s.yaw = lerp(3);
    return s;
  }
  s.yaw = (o.yaw || 0) + (o.spin || 0) * t;
  return s;
}

// How far an object's surface can move between two of its states: the shift of
// its centre plus its turn (summed Euler angles) times its half-diagonal.
function poseTravel(p, o) {
  const shift = Math.hypot(o.x - p.x, o.y - p.y, o.z - p.z);
  const turn = Math.abs((o.rx || 0) - (p.rx || 0)) + Math.abs((o.yaw || 0) - (p.yaw || 0)) + Math.abs((o.rz || 0) - (p.rz || 0));
  const reach = o.sdf ? 0.5 * o.sdf.cell * Math.hypot(o.sdf.dims[0], o.sdf.dims[1], o.sdf.dims[2]) : 0.5;
  return shift + turn * reach;
}
// the fastest object between two samples of the same object list (emitters:
// their stream speed when on); speeds above `lim` are jumps and ask for nothing
function objectSpeed(a, b, dt, lim) {
  let fastest = 0;
  b.forEach((o, i) => {
    if (o.kind === 'emitter') {
      if (o.on) fastest = Math.max(fastest, o.speed || 0);
      return;
    }
    const p = a && a[i];
    if (!p || p.id !== o.id) return;
    const speed = poseTravel(p, o) / dt;
    if (!lim || speed <= lim) fastest = Math.max(fastest, speed);
  });
  return fastest;
}

// smooth value noise in [-1, 1] over an integer lattice hash (the emitters' eddies)
function latticeHash(x, y, z) {
  let h = (x * 374761393 + y * 668265263 + z * 2147483647) | 0;
  h = Math.imul(h ^ (h >>> 13), 1274126177);
  return ((h ^ (h >>> 16)) >>> 0) / 4294967295;
}
function valueNoise(x, y, z) {
  const ix = Math.floor(x), iy = Math.floor(y), iz = Math.floor(z);
  const fade = (f) => f * f * (3 - 2 * f);
  const u = fade(x - ix), v = fade(y - iy), w = fade(z - iz);
  const mix = (a, b, t) => a + (b - a) * t;
  const c = (i, j, k) => latticeHash(ix + i, iy + j, iz + k);
  const lo = mix(mix(c(0, 0, 0), c(1, 0, 0), u), mix(c(0, 1, 0), c(1, 1, 0), u), v);
  const hi = mix(mix(c(0, 0, 1), c(1, 0, 1), u), mix(c(0, 1, 1), c(1, 1, 1), u), v);
  return 2 * mix(lo, hi, w) - 1;
}

// an emitter's frame: the stream direction (azimuth az, elevation el in degrees)
// and two unit vectors across the disc
function emitterFrame(o) {
  const az = o.az * Math.PI / 180, el = o.el * Math.PI / 180;
  const dir = [Math.cos(el) * Math.cos(az), Math.sin(el), Math.cos(el) * Math.sin(az)];
  const up = Math.abs(dir[1]) > 0.95 ? [1, 0, 0] : [0, 1, 0];
  const cross = (p, q) => [p[1] * q[2] - p[2] * q[1], p[2] * q[0] - p[0] * q[2], p[0] * q[1] - p[1] * q[0]];
  let across = cross(dir, up);
  const l = Math.hypot(across[0], across[1], across[2]) || 1;
  across = across.map((c) => c / l);
  return { dir, across, across2: cross(dir, across) };
}

class SceneRunner {
  constructor(sim, rand) {
    this.sim = sim;
    this.prev = new Map();   // object id -> last pose (x, y, z, R)
    this.em = new Map();     // emitter id -> release state
    this.cols = [];
    this.rand = rand || Math.random;
  }
  // collider record of a mesh object; velocity from the pose change over dt and
  // angular velocity from dR = R Rprev^T ~ I + [w]x dt, both dropped above colMax
  collider(o, dt, lim2) {
    const R = eulerRows(o.rx || 0, o.yaw || 0, o.rz || 0), p = this.prev.get(o.id);
    this.prev.set(o.id, { x: o.x, y: o.y, z: o.z, R });
    const c = { sdf: o.sdf, x: o.x, y: o.y, z: o.z, R, vx: 0, vy: 0, vz: 0, wx: 0, wy: 0, wz: 0 };
    if (!p) return c;
    const v = [(o.x - p.x) / dt, (o.y - p.y) / dt, (o.z - p.z) / dt];
    if (v[0] * v[0] + v[1] * v[1] + v[2] * v[2] < lim2) { c.vx = v[0]; c.vy = v[1]; c.vz = v[2]; }
    const dR = (r, q) => R[r * 3] * p.R[q * 3] + R[r * 3 + 1] * p.R[q * 3 + 1] + R[r * 3 + 2] * p.R[q * 3 + 2];
    const w = [(dR(2, 1) - dR(1, 2)) / (2 * dt), (dR(0, 2) - dR(2, 0)) / (2 * dt), (dR(1, 0) - dR(0, 1)) / (2 * dt)];
    if (w[0] * w[0] + w[1] * w[1] + w[2] * w[2] < lim2) { c.wx = w[0]; c.wy = w[1]; c.wz = w[2]; }
    return c;
  }
  step(dt, states, par) {
    const lim2 = Math.pow(par.colMax || 20, 2);
    this.scale = par.scale || 1;
    const cols = [];
    for (const o of states) {
      if (o.kind === 'mesh') { if (o.sdf && !o.water) cols.push(this.collider(o, dt, lim2)); }
      else if (o.on) this.emit(o, dt);
    }
    this.cols = cols;
    return this.sim.step(dt, Object.assign({}, par, { cols }));
  }
  // An emitter: a disc of radius o.radius at (x, y, z) releasing water along its
  // direction at o.speed (scene units: divided by the scene scale), with a
  // velocity spread, and eddies of strength o.turb carried along the stream.
  // mode 'random': a stream of single particles at the disc's volume flux;
  // otherwise a square lattice of the disc each time the stream has advanced one
  // spacing.
  emit(o, dt) {
    const sim = this.sim, d = sim.d, rnd = this.rand, F = emitterFrame(o), [dx, dy, dz] = F.dir;
    let st = this.em.get(o.id);
    if (!st) { st = { phase: 0, acc: 0, key: '', pts: null, t: 0, dist: 0 }; this.em.set(o.id, st); }
    const ks = 1 / (this.scale || 1), spread = (o.spread || 0) * ks, v = o.speed * ks, flow = Math.max(v, 0.3 * ks), turb = o.turb || 0;
    st.t += dt; st.dist += flow * dt;
    // the point at disc coordinates (a, c), `adv` along the stream
    const at = (a, c, adv, q) => [o.x, o.y, o.z][q] + F.across[q] * a + F.across2[q] * c + F.dir[q] * adv;
    // eddies: noise fields fixed in the travelling stream, large across the disc plus smaller ones, slowly changing
    const eddy = (a, c, adv) => {
      if (turb <= 0) return [0, 0, 0];
      const r = Math.max(o.radius, 0.08), k1 = 1.4 / r, k2 = 4.2 / r, zz = (st.dist - adv) / (1.6 * r + 0.25) + st.t * 0.5;
      const field = (ox, oy, oz) => valueNoise(a * k1 + ox, c * k1 + oy, zz + oz) * 0.7 + valueNoise(a * k2 + oz, c * k2 + ox, zz * 2.3 + oy) * 0.3;
      const tv = turb * Math.max(v, 0.6 * ks), along = field(0, 0, 0) * 0.5;
      const s1 = field(31.7, 5.2, 9.1) * 0.9 * tv, s2 = field(-12.4, 77.3, 3.3) * 0.9 * tv;
      const gain = 1 + 0.6 * turb * field(53.1, -8.8, 21.7), jit = turb * 0.1 * Math.max(v, 0.6 * ks);
      void along;
      const out = [];
      for (let q = 0; q < 3; q++) out.push(F.dir[q] * v * (gain - 1) + F.across[q] * s1 + F.across2[q] * s2 + (rnd() - 0.5) * 2 * jit);
      return out;
    };
    if (o.mode === 'random') {
      st.acc += Math.PI * o.radius * o.radius * flow / (d * d * d) * (o.density || 1) * dt;
      const n = Math.floor(st.acc);
      st.acc -= n;
      for (let k = 0; k < n; k++) {
        const ang = rnd() * Math.PI * 2, rr = o.radius * Math.sqrt(rnd());
        const a = Math.cos(ang) * rr, c = Math.sin(ang) * rr, adv = rnd() * flow * dt, e = eddy(a, c, adv);
        const jx = (rnd() - 0.5) * 2 * spread, jy = (rnd() - 0.5) * 2 * spread, jz = (rnd() - 0.5) * 2 * spread;
        if (!sim.add(at(a, c, adv, 0), at(a, c, adv, 1), at(a, c, adv, 2), dx * v + jx + e[0], dy * v + jy + e[1], dz * v + jz + e[2])) break;
      }
      return;
    }
    const key = o.radius + '|' + d;
    if (st.key !== key) {   // the disc's lattice points at half-spacing offsets (one point for a tiny disc)
      const pts = [], R = Math.ceil(o.radius / d) + 1;
      for (let i = -R; i <= R; i++)
        for (let j = -R; j <= R; j++) {
          const a = (i + 0.5) * d, c = (j + 0.5) * d;
          if (a * a + c * c <= o.radius * o.radius) pts.push(a, c);
        }
      st.pts = pts.length ? pts : [0, 0];
      st.key = key;
    }
    st.phase += flow * dt;
    for (let layer = 0; st.phase >= d && layer < 8; layer++) {
      st.phase -= d;
      const adv = st.phase;
      for (let k = 0; k < st.pts.length; k += 2) {
        const a = st.pts[k], c = st.pts[k + 1], e = eddy(a, c, adv);
        const jx = spread ? (rnd() - 0.5) * 2 * spread : 0, jy = spread ? (rnd() - 0.5) * 2 * spread : 0;
        const jz = spread ? (rnd() - 0.5) * 2 * spread : 0;
        if (!sim.add(at(a, c, adv, 0), at(a, c, adv, 1), at(a, c, adv, 2), dx * v + jx + e[0], dy * v + jy + e[1], dz * v + jz + e[2])) return;
      }
    }
  }
}

// Per-frame numbers of a bake (bakeDiag, src/flip/bake.js:60-79) from a readback
// (x, y, z, speed + 64 * foam level; dead slots at y < -1e5): living, gone and
// non-finite points, points outside the box or at the speed cap, mean and
// largest speed, mean foam and foamy points, centre, bounds and a 16-bin height
// histogram.
function bakeDiag(dst, n, b, vcap) {
  const H = 16, hist = new Array(H).fill(0), binH = (b.y1 - b.y0) / H, cap = Math.min(vcap || 14, 63) * 0.98, e = 1e-3;
  const lo = [Infinity, Infinity, Infinity], hi = [-Infinity, -Infinity, -Infinity], sum = [0, 0, 0];
  let alive = 0, bad = 0, out = 0, capped = 0, foamy = 0, speedSum = 0, smax = 0, levels = 0;
  for (let i = 0; i < n; i++) {
    const p = [dst[i * 4], dst[i * 4 + 1], dst[i * 4 + 2]], w = dst[i * 4 + 3];
    if (!(Number.isFinite(p[0]) && Number.isFinite(p[1]) && Number.isFinite(p[2]) && Number.isFinite(w))) { bad++; continue; }
    if (p[1] < -1e5) continue;
    alive++;
    const lv = Math.floor(w / 64), sp = w - lv * 64;
    for (let a = 0; a < 3; a++) sum[a] += p[a];
    speedSum += sp; if (sp > smax) smax = sp;
    levels += lv; if (lv >= 8) foamy++;
    if (sp >= cap) capped++;
    if (p[0] < b.x0 - e || p[0] > b.x1 + e || p[1] < b.y0 - e || p[1] > b.y1 + e || p[2] < b.z0 - e || p[2] > b.z1 + e) out++;
    for (let a = 0; a < 3; a++) { if (p[a] < lo[a]) lo[a] = p[a]; if (p[a] > hi[a]) hi[a] = p[a]; }
    hist[Math.max(0, Math.min(H - 1, Math.floor((p[1] - b.y0) / binH)))]++;
  }
  const a = Math.max(alive, 1);
  return { alive, gone: n - alive - bad, nonFinite: bad, outsideBox: out, atWaterSpeedLimit: capped, meanSpeed: speedSum / a,
    maxSpeed: smax, meanFoam: levels / a / 15, foamyPoints: foamy, center: [sum[0] / a, sum[1] / a, sum[2] / a],
    bbox: alive ? [lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]] : null, heightHist: hist };
}

// The bake loop: desc = { d, maxN, size, open, cellRatio (or ratio), fills:
// [{t:'box', x0..z1} | {t:'sphere', x, y, z, r, vy} | {t:'vol', sdf, x, y, z, R,
// lo, hi}], fps, frames, substeps (least per frame), maxSub, cfl, g, iters, flip,
// comp, wind, drag, foamTau, gust, foamSens, colMax, vmaxW, scale, tear, spr,
// sprayCap, air, drip, nb, ocean, nbLook, fastAir, sortP, objs, events ([{t, x, y,
// z, r}] balls of water dropped at time t) }.  onFrame(frame, info) gets every
// frame as a cache frame (packFrame; with whitewater and deep-water parts when
// they are on) and {f, F, n, elapsed, full, sub, diag}.  Resolves true when all
// frames are done.
async function bake(desc, device, onFrame, isCancelled, rand) {
  const { FlipSim } = require('./sim.js');
  const sim = new FlipSim(device, { d: desc.d, maxN: desc.maxN, size: desc.size, open: desc.open | 0, ratio: desc.cellRatio || desc.ratio });
  await sim.whenReady;
  sim.recycle = false; sim.cap = desc.maxN; sim.opt.fastAir = !!desc.fastAir; sim.opt.sort = !!desc.sortP;
  if (rand) sim.rand = rand;
  if (desc.spr && desc.sprayCap) sim.setSprayCap(desc.sprayCap);
  if (desc.nbLook) sim.setDeepLook(true);
  const runner = new SceneRunner(sim, rand);
  const oc = desc.ocean && desc.ocean.on ? desc.ocean : null, nbc = desc.nb && (desc.nb.on || oc) ? desc.nb : null;
  const fills = desc.fills || [];
  if (nbc || oc) nbSeed(sim, fills, nbc || { on: true, band: 3 }, oc, 0, desc.g, 1);
  else {
    for (const f of fills) {
      if (f.t === 'box') sim.fillBox(f.x0, f.x1, f.y0, f.y1, f.z0, f.z1);
      else if (f.t === 'vol') fillVolume(sim, f);
      else sim.fillSphere(f.x, f.y, f.z, f.r, 0, f.vy || 0, 0);
    }
  }
  const dst = new Float32Array(Math.max(desc.maxN, 1) * 4);
  const F = desc.frames, dtF = 1 / desc.fps, minSub = desc.substeps || 1, cfl = desc.cfl || 1.5, maxSub = Math.max(minSub, desc.maxSub || 16);
  const events = (desc.events || []).slice().sort((a, b) => a.t - b.t);
  const objs = desc.objs || [];
  const statesAt = (t) => objs.map((o) => sampleState(o, t));
  const stepSettings = (t) => ({ g: desc.g, iters: desc.iters, flip: desc.flip, comp: desc.comp, wind: desc.wind, drag: desc.drag,
    foamTau: desc.foamTau, gust: desc.gust, foamSens: desc.foamSens, colMax: desc.colMax, vmax: desc.vmaxW, scale: desc.scale || 1,
    vs: desc.scale || 1, tear: desc.tear || 0, spr: desc.spr || null, nb: nbc, ocean: oc, time: t, air: desc.air || null,
    drip: desc.drip || null });
  const t0 = Date.now();
  let ev = 0, vmax = 2;
  for (const o of statesAt(0)) if (o.kind === 'emitter') vmax = Math.max(vmax, o.speed || 0);
  if (sim.N > 0) await sim.prime({ g: desc.g, iters: desc.iters, cols: [] });
  let n = await sim.readbackInto(dst);
  onFrame(packFrame(dst, n, sim.b, statesAt(0), desc.scale),
    { f: 0, F, n, elapsed: 0, full: false, diag: Object.assign(bakeDiag(dst, n, sim.b, desc.vmaxW), { sub: 0, blockPool: sim.scap }) });
  let ok = true;
  for (let f = 1; f <= F; f++) {
    // the objects' own motion during this frame is known ahead
    const objV = objectSpeed(statesAt((f - 1) * dtF), statesAt(f * dtF), dtF, desc.colMax);
    const sub = Math.max(minSub, Math.min(maxSub, Math.ceil(dtF * Math.max(vmax, objV) / (cfl * sim.h))));
    const dt = dtF / sub, tf0 = Date.now(), ovf0 = sim.ovfTotal || 0;
    for (let s = 0; s < sub; s++) {
      const t = (f - 1) * dtF + (s + 1) * dt;
      for (; ev < events.length && events[ev].t <= t; ev++) sim.fillSphere(events[ev].x, events[ev].y, events[ev].z, events[ev].r, 0, -1, 0);
      runner.step(dt, statesAt(t), stepSettings(t));
    }
    n = await sim.readbackInto(dst);
    // the next frame's substeps follow this frame's fastest particle (w below 64 is the speed), the solver's
    // own maximum and the emitters
    let fastest = 0;
    for (let i = 0; i < n; i++) fastest = Math.max(fastest, dst[i * 4 + 3] % 64);
    const st = sim.stats || {};
    if (st.vmax > fastest) fastest = st.vmax;
    for (const o of statesAt(f * dtF)) if (o.kind === 'emitter' && o.on) fastest = Math.max(fastest, o.speed || 0);
    vmax = Math.max(1, fastest);
    const diag = Object.assign(bakeDiag(dst, n, sim.b, desc.vmaxW), { sub, objectSpeed: objV, fastestParticle: st.vmax, fluidCells: st.fluid,
      blocksUsed: st.sc, blockPool: sim.scap, blocksNoRoom: (sim.ovfTotal || 0) - ovf0, ms: Date.now() - tf0 });
    const frame = packFrame(dst, n, sim.b, statesAt(f * dtF), desc.scale);
    if (sim.sprayCap) {
      const sd = await sim.readSpray();
      if (sd) {
        frame.sp = packSprayFrame(sd, sim.b, 200000, (desc.spr && desc.spr.ww && desc.spr.ww.foamLife) || 3);
        diag.sprayPoints = frame.sp.alive;
      }
    }
    if (desc.nbLook) {
      const dd = await sim.readDeep();
      if (dd && dd.n) { frame.dp = packDeepFrame(dd.p, dd.n, sim.b, desc.scale); diag.deepPoints = dd.n; }
    }
    onFrame(frame, { f, F, n, elapsed: (Date.now() - t0) / 1000, full: sim.N >= sim.cap, sub, diag });
    if (isCancelled && isCancelled()) { ok = false; break; }
    await new Promise((r) => setTimeout(r, 0));
  }
  sim.destroy();
  return ok;
}

module.exports = { SceneRunner, bake, bakeDiag, keyPose, sampleState, objectSpeed, eulerRows, valueNoise };

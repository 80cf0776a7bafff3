'use strict';
// FlipSim on one MI355X: the JS API of the reference's WebGPU FlipSim
// (src/flip/sim.js:46-547) over the HIP library, through the N-API addon
// (splashtank_hip.node -> libsplashtank_hip.so, the C ABI of include/splashtank.h).
//
// Same constructor options, fields, methods and stage names as the reference, so
// its callers (SceneRunner.step, the bake loop, the benchmark) drive this class
// unchanged.  What differs, and why:
//  - `device` is a HIP device handle ({ ordinal }) from initHip(), not a GPUDevice;
//  - there is no pipeline compile step: `ready` is true and `whenReady` resolves at once;
//  - step() enqueues on the sim's HIP stream; readbackInto, readSpray, readDeep and
//    readStats enqueue their copy at once and return a Promise that a worker
//    thread settles when the copy has landed (the reference's mapAsync), so the
//    event loop runs meanwhile and the result is the state at the call;
//  - `stats` is the last finished readStats() result, started after each step
//    while the pool can grow (the reference's lags one step the same way); read
//    before any has finished, it reads the control words at once;
//  - the block pool defaults to every tile of the box (no growth needed on a
//    288 GB GPU); opt.slots still sets a smaller pool, with the reference's
//    overflow counting and growth policy.
// Whitewater (setSprayCap/readSpray), drips, the live view (pump/outFrame) and the
// deep water points of the narrow band (setDeepLook/readDeep) run on the GPU too,
// as does the ocean side boundary (o.ocean with o.time).

const path = require('path');
const { dripSamples } = require('./sdf.js');

let addon = null;
// The packed live view as a cache-format frame: typed-array views straight into
// the packed words at the offsets of its layout (no copy).  Header word 0 counts
// the living whitewater points, word 1 the deep points; the water sample has
// layout.nW points.
function liveFrame(words, L, b) {
  const buf = words.buffer, at = (wordOff) => words.byteOffset + wordOff * 4;
  const nW = L.nW, nS = Math.min(words[0], L.nS), nD = Math.min(words[1], L.nD);
  const fr = { n: nW, pos: new Uint16Array(buf, at(L.oWp), nW * 3), spd: new Uint8Array(buf, at(L.oWs), nW),
    fm: new Uint8Array(buf, at(L.oWf), nW), b, bytes: words.byteLength };
  if (nS) fr.sp = { n: nS, pos: new Uint16Array(buf, at(L.oSp), nS * 3), pk: new Uint32Array(buf, at(L.oSk), nS), alive: nS };
  if (nD) fr.dp = { n: nD, pos: new Uint16Array(buf, at(L.oDp), nD * 3), spd: new Uint8Array(buf, at(L.oDs), nD) };
  return fr;
}
function loadAddon() {
  if (!addon) addon = require(path.join(__dirname, 'splashtank_hip.node'));
  return addon;
}

const STAGE_IDS = { sort: 0, blocks: 1, clear: 2, p2g: 3, 'narrow band': 4, classify: 5, normFaces: 6, divergence: 7,
  pressure: 8, project: 9, g2p: 10, spray: 11, reseed: 12 };
const DUMP = { tiles: 0, cnt: 1, ctype: 2, pres: 3, rhs: 4, accv: 5, accw: 6, facev: 7, faceo: 8, velg: 9, pos: 10, vel: 11, tmap: 12 };

// initGpu's counterpart (gpu.js:492-509): the HIP device the sims run on
function initHip(ordinal) {
  const a = loadAddon();
  return { ordinal: ordinal | 0, version: a.version(), limits: { maxStorageBufferBindingSize: Infinity } };
}

// packColliders (gpu.js:42-70): the collider records (32 floats each: grid
// offset and dims, grid origin and cell, the rows of R^T, position, velocity,
// angular velocity; the layout the kernels read) into sim.cf, and the distance
// grids, one copy per sdf.id in first-use order, uploaded again only when that
// list of ids changes
function packColliders(sim, cols) {
  const nc = Math.min(cols.length, 32), cf = sim.cf;
  const grids = new Map();   // sdf.id -> {offset, data}
  let words = 0;
  for (const c of cols.slice(0, nc)) {
    if (grids.has(c.sdf.id)) continue;
    grids.set(c.sdf.id, { offset: words, data: c.sdf.data });
    words += c.sdf.data.length;
  }
  const sig = Array.from(grids.keys()).join(',') + ',';
  if (sig !== sim.sdfSig) {
    const all = new Float32Array(Math.max(words, 1));
    for (const g of grids.values()) all.set(g.data, g.offset);
    sim.sdfAll = all.subarray(0, words); sim.sdfSig = sig; sim.sdfSigNum = (sim.sdfSigNum || 0) + 1;
  }
  cols.slice(0, nc).forEach((c, k) => {
    const sd = c.sdf, R = c.R, rec = new Float32Array(32);
    rec.set([grids.get(sd.id).offset, sd.dims[0], sd.dims[1], sd.dims[2]], 0);
    rec.set([sd.origin[0], sd.origin[1], sd.origin[2], sd.cell], 4);
    for (let col = 0; col < 3; col++) rec.set([R[col], R[3 + col], R[6 + col]], 8 + 4 * col);   // column col of R = row col of R^T
    rec.set([c.x, c.y, c.z], 20);
    rec.set([c.vx, c.vy, c.vz], 24);
    rec.set([c.wx, c.wy, c.wz], 28);
    cf.set(rec, k * 32);
  });
  loadAddon().setColliders(sim.handle, cf, nc, sim.sdfAll || new Float32Array(0), sim.sdfSigNum || 0);
  return nc;
}

class FlipSim {
  constructor(device, opt) {
    const a = loadAddon();
    const ratio = Math.max(1.4, Math.min(4, opt.ratio || 2));
    const sz = opt.size || [3, 2.6, 3];
    this.dev = device || { ordinal: 0 };
    this.handle = a.create({ d: opt.d, ratio, size: sz, open: opt.open | 0, maxN: opt.maxN, slots: opt.slots | 0,
      device: (device && device.ordinal) | 0 });
    const D = a.getDims(this.handle);
    this.d = opt.d; this.ratio = ratio; this.h = ratio * opt.d; this.rest = ratio * ratio * ratio;
    this.opt = { noFoam: false, noCull: false, noSolidCache: false, fastAir: false, sort: false, sortEvery: 16, iters: 0 };
    this.nx = D.nx; this.ny = D.ny; this.nz = D.nz;
    const W = this.nx * this.h, H = this.ny * this.h, Dd = this.nz * this.h;
    this.b = { x0: -W / 2, x1: W / 2, y0: 0, y1: H, z0: -Dd / 2, z1: Dd / 2 };
    this.size = [W, H, Dd];
    this.open = opt.open | 0;
    this.MAXN = opt.maxN; this.cap = opt.maxN; this.N = 0; this.ring = 0; this.recycle = true;
    this.sprayCap = 0;
    this.lastMs = 0; this.slowMs = 0; this.inflight = 0; this.out = opt.out || null; this.outN = 0;
    this.ncell = D.ncell; this.tx = D.tx; this.ty = D.ty; this.tz = D.tz; this.ntiles = D.ntiles;
    this.scap = D.scap; this.scapMax = D.scapMax;
    this.pS = new Float32Array(this.MAXN * 4); this.vS = new Float32Array(this.MAXN * 4);
    this.runStart = 0; this.runLen = 0; this.devN = 0;
    this.cf = new Float32Array(32 * 32); this.sdfSig = null;
    this.tm = 0; this.seed = 0; this.nbOn = false; this.sortedEver = true;
    this.nbInit = 0; this.nbFills = []; this.nbFillVols = []; this.nbCfg = null; this.oceanCfg = null;
    this.ready = true; this.whenReady = Promise.resolve(this);
    this.statsCache = null; this.peak = 0; this.grew = 0; this.ovfTotal = 0;
    this.rand = Math.random;   // jitter source of fillBox / fillSphere
  }
  // ---- particles (GpuSim.setCap/clear/add/flushRun, gpu.js:388-408) ----
  // the most particles the sim keeps (at most maxN); a smaller cap drops the ones past it
  setCap(c) {
    this.cap = Math.min(c, this.MAXN);
    this.N = Math.min(this.N, this.cap);
    if (this.ring >= this.cap) this.ring = 0;
  }
  // One particle into host staging: the next free slot while there is room,
  // else (recycle on) the oldest slot of the ring, else refused (false).  w of
  // the position is the speed, w of the velocity (foam) starts at 0.  Slots
  // written one after another form a run that flushRun uploads in one copy.
  add(x, y, z, vx, vy, vz) {
    let slot;
    if (this.N < this.cap) slot = this.N++;
    else if (this.recycle) { slot = this.ring; this.ring = (this.ring + 1 >= this.cap) ? 0 : this.ring + 1; }
    else return false;
    this.pS.set([x, y, z, Math.sqrt(vx * vx + vy * vy + vz * vz)], slot * 4);
    this.vS.set([vx, vy, vz, 0], slot * 4);
    if (this.runLen > 0 && slot === this.runStart + this.runLen) this.runLen++;
    else { this.flushRun(); this.runStart = slot; this.runLen = 1; }
    return true;
  }
  flushRun() {
    if (!this.runLen) return;
    const a = loadAddon(), s = this.runStart * 4, e = (this.runStart + this.runLen) * 4;
    a.upload(this.handle, this.runStart, this.pS.subarray(s, e), this.vS.subarray(s, e), this.runLen);
    this.runLen = 0;
    if (this.devN !== this.N) { a.setCount(this.handle, this.N); this.devN = this.N; }
  }
  // a lattice at spacing d from d/2 in (JS double steps, y outer, x, z inner), each
  // point jittered by up to 0.005 d / 0.12 on each axis (shared.js:285-297)
  jittered(x, y, z, vx, vy, vz) {
    const j = 0.01 * this.d / 0.12, r = this.rand;
    const jx = (r() - 0.5) * j, jy = (r() - 0.5) * j, jz = (r() - 0.5) * j;
    return this.add(x + jx, y + jy, z + jz, vx, vy, vz);
  }
  fillBox(x0, x1, y0, y1, z0, z1) {
    const d = this.d;
    for (let y = y0 + d / 2; y < y1; y += d)
      for (let x = x0 + d / 2; x < x1; x += d)
        for (let z = z0 + d / 2; z < z1; z += d) this.jittered(x, y, z, 0, 0, 0);
  }
  // the lattice of the cube [-R, R]^3 (steps of d from -R) kept inside the ball, moved to the centre
  fillSphere(cx, cy, cz, R, vx, vy, vz) {
    const d = this.d;
    for (let y = -R; y <= R; y += d)
      for (let x = -R; x <= R; x += d)
        for (let z = -R; z <= R; z += d)
          if (x * x + y * y + z * z <= R * R) this.jittered(cx + x, cy + y, cz + z, vx, vy, vz);
  }
  clear() {
    this.flushRun(); this.N = 0; this.ring = 0; this.devN = 0;
    loadAddon().clear(this.handle);
    this.tm = 0; this.nbInit = 1; this.nbFills = []; this.nbFillVols = []; this.outFrame = null;
    this.dripSig = '';   // drip samples come back dry
  }
  // ---- stepping (sim.js:179-243) ----
  omega() { return Math.min(1.9, 2 / (1 + Math.sin(Math.PI / Math.max(this.nx, this.ny, this.nz)))); }
  // the ocean of this step: o.ocean when on, else (o.ocean not given) the one a narrow-band start set up
  oceanOf(o) {
    if (o.ocean && o.ocean.on) return o.ocean;
    return o.ocean === undefined && this.oceanCfg && this.oceanCfg.on ? this.oceanCfg : null;
  }
  // its wave list, made again only when the settings or gravity change
  oceanWaves(oc, g) {
    const key = [oc.height, oc.length, oc.dir, oc.spread, oc.seed, oc.level, g].join('|');
    if (key !== this.oceanKey) {
      this.oceanKey = key;
      this.oceanW = loadAddon().oceanWaves(oc, g, oc.level - this.b.y0);
    }
    return this.oceanW;
  }
  params(dt, o) {
    const nbc = o.nb !== undefined ? o.nb : this.nbCfg;
    const oc = this.oceanOf(o);
    const nbOn = !!(nbc && nbc.on) || !!oc;   // the ocean runs with the narrow band
    const init = nbOn && this.nbInit ? this.nbInit : 0;
    const fl = init ? this.nbFills.slice(0, 8) : [];
    const fills = new Float32Array(64);
    fl.forEach((f, k) => { const q = k * 8; if (f.t === 'box') fills.set([1, f.x0, f.y0, f.z0, f.x1, f.y1, f.z1, 0], q); else fills.set([2, f.x, f.y, f.z, f.r, 0, 0, 0], q); });
    return { dt, g: o.g, flip: o.flip === undefined ? 0.95 : o.flip, comp: o.comp === undefined ? 6 : o.comp,
      omega: o.omega || this.omega(), vmax: o.vmax || 14, tear: o.tear || 0,
      iters: Math.max(1, (this.opt.iters || o.iters) | 0), wind: o.wind || [0, 0, 0], drag: o.drag || 0,
      foamTau: o.foamTau || 1.2, foamSens: o.foamSens || 1, gust: o.gust || 0,
      noFoam: this.opt.noFoam, noCull: this.opt.noCull, fastAir: this.opt.fastAir, noSolidCache: this.opt.noSolidCache,
      nbOn, nbBand: Math.max(2, (nbc && nbc.band) || 3), onlyTiles: this.only === 'tiles', nbInit: init, nfill: fl.length,
      nfv: init ? this.nbFillVols.length : 0, fills, sorMode: this.opt.sorMode || 0, ...this.sprayParams(o),
      // opt-in modes with no counterpart in the reference (off unless asked for): o.rk2 moves
      // particles by the midpoint rule through the grid velocity, o.extrap carries face
      // velocities that many layers into the air after the projection
      rk2: o.rk2 ? 1 : 0, extrapLayers: Math.max(0, o.extrap | 0),
      oceanOn: oc ? 1 : 0, oceanTime: o.time || 0, oceanLevel: oc ? oc.level : 0, oceanLayer: oc ? oc.layer : 0,
      waves: oc ? this.oceanWaves(oc, o.g) : new Float32Array(0) };
  }
  // o.spr, o.spr.ww, o.air (sim.js:203-206, 221-230), flattened for the C ABI; the C side applies the
  // reference's defaults and its "on only with room for points" rule
  sprayParams(o) {
    const sp = o.spr || null, ww = (sp && sp.ww) || {}, air = o.air || null;
    const out = { vs: o.vs || 1 };
    if (sp) {
      Object.assign(out, { sprOn: 1, sprRate: sp.rate, sprMin: sp.min, sprLife: sp.life, sprMist: sp.mist, sprDrag: sp.drag, sprMdrag: sp.mdrag,
        wwFoamLife: ww.foamLife || 0, wwRise: ww.rise || 0, wwCrest: ww.crest === undefined ? 1 : ww.crest,
        wwImpact: ww.impact === undefined ? 0.5 : ww.impact });
    }
    if (air && air.on) Object.assign(out, { airOn: 1, airCouple: air.couple, airDecay: air.decay, airBlur: air.blur === undefined ? 0.35 : air.blur });
    const dr = o.drip || null;
    if (sp && dr && dr.on) Object.assign(out, { dripOn: 1, dripRate: dr.rate, dripDry: dr.dry });
    return out;
  }
  // drip samples on the colliders, rebuilt when the set of collider grids changes (dripUpdate, sim.js:353-363)
  dripUpdate(cols) {
    const sig = cols.map((c) => (c.sdf ? c.sdf.id : 'x')).join(',');
    if (sig === this.dripSig) return;
    this.dripSig = sig;
    const parts = []; let n = 0;
    cols.forEach((c, k) => { if (!c.sdf) return; const a = dripSamples(c.sdf, this.h, k); parts.push(a); n += a.length / 8; });
    const all = new Float32Array(n * 8); let o = 0; parts.forEach((a) => { all.set(a, o); o += a.length; });
    loadAddon().setDrips(this.handle, all);
  }
  prepare(dt, o, probe) {
    this.flushRun();
    if (this.N === 0) return null;
    if (probe) return { prm: this.params(dt, o), nc: 0 };   // settings only: nothing is stepped
    if (!this.only) this.maybeGrow();
    const prm = this.params(dt, o);
    if (prm.dripOn && this.sprayCap > 0) this.dripUpdate((o.cols || []).slice(0, (o.cols || []).length));
    const init = prm.nbInit;
    const cols = init && this.nbFillVols.length ? (o.cols || []).concat(this.nbFillVols) : (o.cols || []);
    const nc = Math.min((o.cols || []).length, packColliders(this, cols));
    this.nbOn = prm.nbOn;
    if (init && !this.only) { this.nbInit = 0; this.nbFills = []; this.nbFillVols = []; }
    if (!this.only) { this.tm += dt; this.seed = (this.seed + 1) >>> 0; }
    return { prm, nc };
  }
  step(dt, o) {
    const t = this.prepare(dt, o);
    if (!t) return undefined;
    const t0 = Date.now();
    loadAddon().step(this.handle, t.prm);
    this.afterStep();
    if (!this.only) this.track(t0);
    return t.nc;
  }
  // after enqueueing steps: while the pool can grow, the control words are read without
  // blocking (readStats in the reference's step, sim.js:237) and `stats` keeps the last
  // landed read; with every tile pooled nothing needs them, so `stats` is read when asked
  afterStep() {
    this.stepped = true;
    if (this.scap < this.scapMax && !this.only) this.readStats().catch(() => {});
    else this.statsCache = null;
  }
  track(t0) { const ms = Date.now() - t0; this.lastMs = ms; this.slowMs = Math.max(this.slowMs, ms); }
  // n substeps with the same settings, as n step() calls; graph: each one a HIP graph replay (plain
  // substeps only).  No counterpart in the reference.
  stepN(dt, o, n, graph) {
    const t = this.prepare(dt, o);
    if (!t) return undefined;
    if (!this.only) { this.tm += dt * (n - 1); this.seed = (this.seed + n - 1) >>> 0; }
    loadAddon().stepN(this.handle, t.prm, n | 0, graph ? 1 : 0);
    this.afterStep();
    return t.nc;
  }
  // times the pressure-solve paths from the current state (left unchanged) and keeps the fastest for
  // steps with opt.sorMode = 3: {mode, width, candidates}.  No counterpart in the reference.
  tuneSor(dt, o, reps) {
    const t = this.prepare(dt, o, true);
    if (!t) return null;
    return loadAddon().tuneSor(this.handle, t.prm, reps || 3);
  }
  // a copy of the state the next substep starts from; restore(snapshot) goes back to it
  snapshot() { return loadAddon().snapshotCreate(this.handle); }
  restore(snap) { loadAddon().snapshotRestore(this.handle, snap); }
  dropSnapshot(snap) { loadAddon().snapshotDestroy(snap); }
  // the named stages of one step (sim.js:245-310): fn() enqueues that stage
  stages() {
    const a = loadAddon(), h = this.handle;
    const names = ['blocks', 'clear', 'p2g'].concat(this.nbOn ? ['narrow band'] : [],
      ['classify', 'normFaces', 'divergence', 'pressure', 'project', 'g2p'], this.sprayCap > 0 ? ['spray'] : [], this.nbOn ? ['reseed'] : []);
    if (this.only === 'tiles') return [{ name: 'blocks', fn: () => a.stage(h, STAGE_IDS.blocks) }];
    return names.map((name) => ({ name, fn: () => a.stage(h, STAGE_IDS[name]) }));
  }
  // one real step with every stage timed on its own (sim.js:312-331)
  async stepTimed(dt, o) {
    const t = this.prepare(dt, o);
    if (!t) return null;
    const a = loadAddon();
    a.beginStep(this.handle, t.prm);
    const out = {};
    for (const st of this.stages()) {
      const t0 = process.hrtime.bigint();
      st.fn(); a.sync(this.handle);
      out[st.name] = Number(process.hrtime.bigint() - t0) / 1e6;
    }
    this.statsCache = null;
    return out;
  }
  // ---- pool (sim.js:137-178) ----
  statsFrom(a) {
    this.statsCache = { sc: a[4], ovf: a[5], fluid: a[0] + a[1], vmax: new Float32Array(a.buffer)[6], seen: false };
    this.peak = Math.max(this.peak, a[4]);
    return this.statsCache;
  }
  get stats() {
    if (this.statsCache) return this.statsCache;
    return this.statsFrom(loadAddon().readCtl(this.handle));
  }
  // the control words of the steps enqueued so far, read without blocking (sim.js:160-168):
  // sets `stats` when they land; one read at a time
  readStats() {
    if (this.statBusy || !this.handle) return this.statPromise;
    this.statBusy = true;
    this.statPromise = loadAddon().readAsync(this.handle, 1, null, 16).then((a) => {
      this.statsFrom(a); this.statBusy = false; return this.statsCache;
    }, (e) => { this.statBusy = false; throw e; });
    return this.statPromise;
  }
  grow(n) {
    n = Math.min(this.scapMax, Math.max(n, this.scap + 1)); if (n <= this.scap) return false;
    loadAddon().grow(this.handle, n);
    this.scap = loadAddon().getDims(this.handle).scap;
    this.grew = (this.grew || 0) + 1;
    return true;
  }
  maybeGrow() {
    // a pool that holds every tile of the box never overflows: no stats read, no sync
    if (this.scap >= this.scapMax || !this.stepped) return;
    const s = this.statsCache; if (!s || s.seen) return;
    s.seen = true; this.ovfTotal += s.ovf;
    const want = s.sc + s.ovf;
    if (s.ovf > 0 || want > 0.8 * this.scap) this.grow(Math.max(Math.ceil(this.scap * 1.5), Math.ceil(want * 1.4) + 8));
  }
  async prime(o) {
    for (let k = 0; k < 8; k++) {
      if (this.statBusy) await this.statPromise;
      this.only = 'tiles'; this.step(1 / 60, o); this.only = null;
      this.statsCache = null;
      const s = await this.readStats(); s.seen = true;
      if (s.ovf === 0 && s.sc <= 0.8 * this.scap) return;
      this.grow(Math.max(Math.ceil(this.scap * 1.5), Math.ceil((s.sc + s.ovf) * 1.4) + 8));
    }
  }
  // ---- readback (sim.js:439-495) ----
  readbackInto(dst) {
    this.flushRun();
    const n = this.N;
    if (n <= 0) return Promise.resolve(n);
    return loadAddon().readAsync(this.handle, 0, dst, n);
  }
  async activeCells() { const s = this.stats; this.blocksUsed = s.sc; return s.fluid; }
  memBytes() { return loadAddon().memBytes(this.handle); }
  // FP64 norms of the last substep's pressure solve and projection (no counterpart in the
  // reference, whose solve reports nothing): {resL2, resMax, rhsL2, divL2, divMax, errL2, errMax, fluid}
  solveNorms() { return loadAddon().solveNorms(this.handle); }
  dump(what) { return loadAddon().dump(this.handle, DUMP[what]); }
  // one column / row of cells for debugging (sim.js:457-482)
  cellsAt(list) {
    const tmap = new Uint32Array(this.dump('tmap')), tiles = new Uint32Array(this.dump('tiles'));
    const rank = new Map(); tiles.forEach((t, q) => rank.set(t, q));
    const ty = new Uint32Array(this.dump('ctype')), cn = new Int32Array(this.dump('cnt')), pr = new Float32Array(this.dump('pres'));
    const vg = this.nbOn ? new Float32Array(this.dump('velg')) : null;
    return list.map(([i, j, k]) => {
      const t = (i >> 3) + this.tx * ((j >> 3) + this.ty * (k >> 3));
      if (tmap[t] === 0xffffffff) return null;
      const c = rank.get(t) * 512 + (i & 7) + 8 * ((j & 7) + 8 * (k & 7));
      return { ty: ty[c] & 255, col: (ty[c] >> 8) & 255, phi: +(8 - (ty[c] >>> 16) / 4096).toFixed(2), n: cn[c],
        vy: vg ? +vg[c * 3 + 1].toFixed(3) : 0, vx: vg ? vg[c * 3] : 0, p: +pr[c].toFixed(4) };
    });
  }
  async dbgColumn(i, k) {
    const js = []; for (let j = 0; j < this.ny; j++) js.push([i, j, k]);
    return this.cellsAt(js).map((c, j) => (c ? Object.assign({ j }, c) : { j, none: true }));
  }
  async dbgRow(j, k) {
    const is = []; for (let i = 0; i < this.nx; i++) is.push([i, j, k]);
    return this.cellsAt(is).map((c) => (c ? c.ty + '/' + c.phi.toFixed(1) + '/' + c.n + '/' + c.vx.toFixed(3) + '/' + c.vy.toFixed(3) : '-'));
  }
  destroy() { if (this.handle) { loadAddon().destroy(this.handle); this.handle = null; } this.dead = true; }
  // ---- outside the substep: whitewater and the live view ----
  // room for n whitewater points, 0 turns the spray stage off (sim.js:333-340)
  setSprayCap(n) { n = Math.max(0, n | 0); loadAddon().setSprayCap(this.handle, n); this.sprayCap = n; }
  // every whitewater slot handed out so far: positions with life in w, velocities with kind in w (sim.js:423-437)
  readSpray() { return this.sprayCap ? loadAddon().readAsync(this.handle, 2, null, this.sprayCap) : Promise.resolve(null); }
  // narrow band look only (sim.js:405-419): deep water cells (water on the grid with no particles) as points
  setDeepLook(on) { this.deepOn = !!on; }
  // the deep water points of the last step (x, y, z, 4096 + speed), in active-cell order, or null
  readDeep() {
    return this.deepOn && this.nbOn && this.handle && this.N ? loadAddon().readAsync(this.handle, 3, null, 0) : Promise.resolve(null);
  }
  // the live view (sim.js:504-525): packed on the GPU (st_pack_view), here read back at once; outFrame is a
  // frame in the cache format (views into the packed words) that the drawing uploads as it is
  pump() {
    if (!this.handle || this.N === 0) return;
    this.flushRun();
    const r = loadAddon().packView(this.handle, this.dispMax || 300000);
    this.outFrame = liveFrame(r.words, r.layout, Object.assign({}, this.b));
    this.outN = r.layout.nW; this.gen = (this.gen || 0) + 1;
  }
}

// The adjoint of the substep (st_adj_*; no counterpart in the reference):
// record(dt, o) runs a substep as step() does and keeps its tape; seed(gPos, gVel)
// sets dL/d(state) (Float32Array, 4 per particle in id order, w ignored);
// backward() reverses the last recorded substep; grad(gPos, gVel) reads
// dL/d(state before it); colGrad(ncols) the 9 floats per collider of that
// substep (dL/d centre, velocity, angular velocity).
class FlipAdjoint {
  constructor(sim, maxSteps) { this.sim = sim; this.handle = loadAddon().adjCreate(sim.handle, maxSteps | 0); }
  record(dt, o) {
    const t = this.sim.prepare(dt, o);
    if (!t) return undefined;
    loadAddon().adjRecord(this.handle, t.prm);
    this.sim.statsCache = null; this.sim.stepped = true;
    return t.nc;
  }
  seed(gPos, gVel) { loadAddon().adjSeed(this.handle, gPos, gVel); }
  backward() { loadAddon().adjBackward(this.handle); }
  grad(gPos, gVel) { loadAddon().adjGrad(this.handle, gPos, gVel); }
  colGrad(ncols) { return loadAddon().adjColGrad(this.handle, ncols | 0); }
  destroy() { if (this.handle) { loadAddon().adjDestroy(this.handle); this.handle = null; } }
}

// Many tanks stepped together, one kernel launch per stage for all of them
// (st_batch_*; the reference steps one FlipSim at a time).  Each tank keeps its
// FlipSim API; step(dt, o) takes one `o` for all tanks or an array, one per tank,
// and gives every tank the result its own step(dt, o) would.
class FlipBatch {
  constructor(sims) {
    this.sims = sims.slice();
    this.handle = loadAddon().batchCreate(this.sims.map((s) => s.handle));
  }
  step(dt, o) {
    const os = Array.isArray(o) ? o : this.sims.map(() => o);
    if (os.length !== this.sims.length) throw new Error('FlipBatch.step: one o per tank');
    const prms = [];
    let any = null;
    this.sims.forEach((s, k) => {
      const t = s.prepare(dt, os[k]);
      prms.push(t ? t.prm : null);
      if (t) { any = t.prm; s.statsCache = null; s.stepped = true; }
    });
    if (!any) return;
    // tanks without particles sit the step out; any valid block stands in for theirs
    loadAddon().batchStep(this.handle, prms.map((p) => p || any));
  }
  // n batched substeps with the same settings; graph: HIP graph replays
  stepN(dt, o, n, graph) {
    const prms = this.prms(dt, o, false);
    if (prms) loadAddon().batchStepN(this.handle, prms, n | 0, graph ? 1 : 0);
    if (prms) this.sims.forEach((s) => { if (!s.only && s.N) { s.tm += dt * (n - 1); s.seed = (s.seed + n - 1) >>> 0; } });
  }
  // times batched substeps at several SOR widths and keeps the fastest (st_batch_tune_sor)
  tuneSor(dt, o, reps) {
    const prms = this.prms(dt, o, true);
    return prms ? loadAddon().batchTuneSor(this.handle, prms, reps || 3) : null;
  }
  prms(dt, o, probe) {
    const os = Array.isArray(o) ? o : this.sims.map(() => o);
    if (os.length !== this.sims.length) throw new Error('FlipBatch: one o per tank');
    let any = null;
    const prms = this.sims.map((s, k) => {
      const t = s.prepare(dt, os[k], probe);
      if (t) { any = t.prm; if (!probe) { s.statsCache = null; s.stepped = true; } }
      return t ? t.prm : null;
    });
    return any ? prms.map((p) => p || any) : null;
  }
  sync() { loadAddon().batchSync(this.handle); }
  destroy() { if (this.handle) { loadAddon().batchDestroy(this.handle); this.handle = null; } }
}

module.exports = { FlipSim, FlipBatch, FlipAdjoint, initHip, packColliders, STAGE_IDS };

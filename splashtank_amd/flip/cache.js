'use strict';
// The bake cache on disk (SPLC), so frames baked on the MI355X play in the
// reference's page: packFrame turns a readback into a cache frame (the
// reference's packFrameBuf, src/gpu.js:525-560), packSprayFrame / packDeepFrame
// the whitewater and deep-water parts (packSprayFrame, src/flip/bake.js:43-58;
// packDeep, src/gpu.js:512-524), writeCache lays frames out as version 2
// (writeCacheParts, src/shared.js:686-709) and readCache reads version 2 and the
// native baker's version 1 (parseCache2 / parseCache, shared.js:710-762;
// native/splash.cpp:343-360).
//
// Version 2 layout (little endian): 'SPLC', u32 2, u32 JSON length, the JSON
// (spacing, fps, object ids, box, scene, settings), zero padding to 4 bytes,
// u32 frame count; per frame u32 points, u32 flags (1 white water levels
// follow, 2 spray follows, 4 deep water points follow), 6 f32 per object (x y z
// rx yaw rz), 3 u16 per point (position across the box, 0..65535), u8 per
// point (speed in 16ths of m/s times the scene scale, 255 = no point), [u8 per
// point white water level * 17], zero padding to 4 bytes; then with flag 2 u32
// count, 3 u16 each, padding, u32 code each; with flag 4 u32 count, 3 u16 and a
// u8 each, padding.

// a coordinate across the box as a 16-bit code (round half up, clamped)
function q16(v, lo, k) {
  const q = Math.round((v - lo) * k);
  return q < 0 ? 0 : (q > 65535 ? 65535 : q);
}
function boxScales(b) {
  return [65535 / (b.x1 - b.x0), 65535 / (b.y1 - b.y0), 65535 / (b.z1 - b.z0)];
}
function putPos(pos, k, x, y, z, b, K) {
  pos[k * 3] = q16(x, b.x0, K[0]);
  pos[k * 3 + 1] = q16(y, b.y0, K[1]);
  pos[k * 3 + 2] = q16(z, b.z0, K[2]);
}
// a speed in 16ths (times the scene scale) as a byte, 191 at most
function speedByte(s) { return s > 191 ? 191 : s; }

function packObjs(states) {
  const objs = new Float32Array((states || []).length * 6);
  (states || []).forEach((s, k) => objs.set([s.x, s.y, s.z, s.rx || 0, s.yaw || 0, s.rz || 0], k * 6));
  return objs;
}

// The 32-bit code of a whitewater point (packSpray, wgsl-whitewater.js:241-253):
// bit 31 set; bits 29-30 the kind (spray 00, mist 10, foam 01, bubble 11, read
// as bit 30 = mist-or-bubble, bit 29 = foam-or-bubble); bits 21-28 the speed in
// quarter units per second (255 at most); below, for spray and mist the flight
// direction as an octahedral map of the unit sphere (10 bits each of u, v), for
// foam and bubbles the share of life left in 10 bits.
const KIND_BITS = [0, 1 << 30, 1 << 29, 3 << 29];
function packSprayCode(vx, vy, vz, kind, lifeFrac) {
  const k = (kind === true ? 1 : kind | 0) & 3;
  const s = Math.hypot(vx, vy, vz);
  const head = (0x80000000 | KIND_BITS[k] | (Math.min(Math.floor(s * 4), 255) << 21)) >>> 0;
  if (k >= 2) return (head | Math.round(Math.min(Math.max(lifeFrac || 0, 0), 1) * 1023)) >>> 0;
  // the direction on the octahedron |x| + |y| + |z| = 1, the lower half folded out over the upper
  const dir = s > 1e-6 ? [vx / s, vy / s, vz / s] : [0, 1, 0];
  const l1 = Math.abs(dir[0]) + Math.abs(dir[1]) + Math.abs(dir[2]);
  let u = dir[0] / l1, v = dir[1] / l1;
  if (dir[2] < 0) {
    const fu = (1 - Math.abs(v)) * (u >= 0 ? 1 : -1), fv = (1 - Math.abs(u)) * (v >= 0 ? 1 : -1);
    u = fu; v = fv;
  }
  const code10 = (t) => Math.min(Math.max(Math.round((t * 0.5 + 0.5) * 1023), 0), 1023);
  return (head | (code10(v) << 10) | code10(u)) >>> 0;
}

// readback (id order: x, y, z, w) -> cache frame.  w is speed + 64 * foam level
// for water; a w of 4096 and more is a deep-water point, a w with its sign bit
// set a whitewater code (the live view's buffers hold both); dead slots have
// y < -1e5.  Water keeps its slot order; empty slots after the last water point
// are dropped, the ones before it get speed byte 255.
function packFrame(buf, n, b, states, vs) {
  const ks = 16 * (vs || 1), U = new Uint32Array(buf.buffer, buf.byteOffset, buf.length), K = boxScales(b);
  const dead = (i) => buf[i * 4 + 1] <= -1e5;
  const isSpray = (i) => U[i * 4 + 3] >= 0x80000000;
  const isDeep = (i) => !isSpray(i) && buf[i * 4 + 3] >= 4096;
  let last = -1, ns = 0, firstDeep = -1;
  for (let i = 0; i < n; i++) {
    if (dead(i)) continue;
    if (isSpray(i)) ns++;
    else if (isDeep(i)) { if (firstDeep < 0) firstDeep = i; }
    else last = i;
  }
  const m = last + 1;
  const pos = new Uint16Array(m * 3), spd = new Uint8Array(m), fm = new Uint8Array(m);
  const sp = ns ? { n: ns, pos: new Uint16Array(ns * 3), pk: new Uint32Array(ns), alive: ns } : null;
  let foam = false, si = 0;
  for (let i = 0; i < n; i++) {
    const water = i < m;
    if (buf[i * 4 + 1] < -1e5 || isDeep(i)) { if (water) spd[i] = 255; continue; }
    if (isSpray(i)) {
      if (water) spd[i] = 255;
      putPos(sp.pos, si, buf[i * 4], buf[i * 4 + 1], buf[i * 4 + 2], b, K);
      sp.pk[si++] = U[i * 4 + 3];
      continue;
    }
    putPos(pos, i, buf[i * 4], buf[i * 4 + 1], buf[i * 4 + 2], b, K);
    const w = buf[i * 4 + 3], lv = Math.floor(w / 64);
    spd[i] = speedByte((w - lv * 64) * ks);
    if (lv) { fm[i] = lv * 17; foam = true; }
  }
  const fr = { n: m, pos, spd, objs: packObjs(states) };
  if (foam) fr.fm = fm;
  if (sp) fr.sp = sp;
  if (firstDeep >= 0) {
    const dp = packDeepFrame(buf.subarray(firstDeep * 4, n * 4), n - firstDeep, b, vs);
    if (dp.n) fr.dp = dp;
  }
  return fr;
}

// deep water points (x, y, z, 4096 + speed) -> positions across the box and speed bytes
function packDeepFrame(p, n, b, vs) {
  const ks = 16 * (vs || 1), K = boxScales(b);
  const pos = new Uint16Array(n * 3), spd = new Uint8Array(n);
  let k = 0;
  for (let i = 0; i < n; i++) {
    if (p[i * 4 + 1] <= -1e5) continue;
    putPos(pos, k, p[i * 4], p[i * 4 + 1], p[i * 4 + 2], b, K);
    spd[k++] = speedByte(Math.max(0, (p[i * 4 + 3] - 4096) * ks));
  }
  return { n: k, pos: k < n ? pos.slice(0, k * 3) : pos, spd: k < n ? spd.slice(0, k) : spd };
}

// living whitewater slots of readSpray ({n, p, v}: life in p.w, kind in v.w) as a
// frame part: every step-th slot, step = ceil(living / cap), so a point keeps
// its place in the sample while it lives
function packSprayFrame(sd, b, cap, foamLife) {
  const life = foamLife || 3, K = boxScales(b);
  const living = (j) => sd.p[j * 4 + 3] > 0;
  let alive = 0;
  for (let j = 0; j < sd.n; j++) if (living(j)) alive++;
  const step = Math.max(1, Math.ceil(alive / cap));
  const picked = [];
  for (let j = 0; j < sd.n; j += step) if (living(j)) picked.push(j);
  const pos = new Uint16Array(picked.length * 3), pk = new Uint32Array(picked.length);
  picked.forEach((j, k) => {
    putPos(pos, k, sd.p[j * 4], sd.p[j * 4 + 1], sd.p[j * 4 + 2], b, K);
    pk[k] = packSprayCode(sd.v[j * 4], sd.v[j * 4 + 1], sd.v[j * 4 + 2], Math.round(sd.v[j * 4 + 3]), sd.p[j * 4 + 3] / life);
  });
  return { n: picked.length, pos, pk, alive };
}

// meta: { d, fps, ids, b, org, maxN, seconds, engine, scene, desc, diag, stableIds, scale, spdMs }
function writeCache(meta, frames) {
  const m = { d: meta.d, fps: meta.fps, ids: meta.ids || [], b: meta.b, org: meta.org, maxN: meta.maxN, seconds: meta.seconds,
    engine: meta.engine, scene: meta.scene, desc: meta.desc, diag: meta.diag, stableIds: !!meta.stableIds, scale: meta.scale || 1,
    spdMs: !!meta.spdMs };
  const js = Buffer.from(JSON.stringify(m), 'utf8');
  const zeros = (k) => Buffer.alloc((4 - (k % 4)) % 4);
  const u32 = (...v) => Buffer.from(new Uint32Array(v).buffer);
  const raw = (a) => Buffer.from(a.buffer, a.byteOffset, a.byteLength);
  const parts = [Buffer.from('SPLC', 'latin1'), u32(2, js.length), js, zeros(js.length), u32(frames.length)];
  for (const fr of frames) {
    const sp = fr.sp && fr.sp.n ? fr.sp : null, dp = fr.dp && fr.dp.n ? fr.dp : null;
    const flags = (fr.fm ? 1 : 0) | (sp ? 2 : 0) | (dp ? 4 : 0);
    parts.push(u32(fr.n, flags), raw(fr.objs), raw(fr.pos), raw(fr.spd));
    if (fr.fm) parts.push(raw(fr.fm));
    parts.push(zeros(fr.n * (fr.fm ? 8 : 7)));
    if (sp) parts.push(u32(sp.n), raw(sp.pos), zeros(sp.n * 6), raw(sp.pk));
    if (dp) parts.push(u32(dp.n), raw(dp.pos), raw(dp.spd), zeros(dp.n * 7));
  }
  return Buffer.concat(parts);
}

function readCache(buf) {
  const dv = new DataView(buf.buffer, buf.byteOffset, buf.byteLength), len = buf.byteLength;
  if (len < 24 || buf.toString('latin1', 0, 4) !== 'SPLC') throw new Error('not an SPLC cache');
  const slice = (o, n, T) => new T(buf.buffer.slice(buf.byteOffset + o, buf.byteOffset + o + n * T.BYTES_PER_ELEMENT));
  const pad = (k) => (4 - (k % 4)) % 4;
  const frames = [];
  let maxN = 0;
  if (dv.getUint32(4, true) === 2) {
    let o = 8;
    const ml = dv.getUint32(o, true); o += 4;
    const meta = JSON.parse(buf.toString('utf8', o, o + ml)); o += ml + pad(ml);
    const nf = dv.getUint32(o, true); o += 4;
    const nobj = (meta.ids || []).length;
    for (let f = 0; f < nf && o + 8 <= len; f++) {
      const n = dv.getUint32(o, true), fl = dv.getUint32(o + 4, true); o += 8;
      const per = (fl & 1) ? 8 : 7;
      if (o + nobj * 24 + n * per > len) break;
      const fr = { n, objs: slice(o, nobj * 6, Float32Array) }; o += nobj * 24;
      fr.pos = slice(o, n * 3, Uint16Array); o += n * 6;
      fr.spd = slice(o, n, Uint8Array); o += n;
      if (fl & 1) { fr.fm = slice(o, n, Uint8Array); o += n; }
      o += pad(n * per);
      if (fl & 2) {
        const ns = dv.getUint32(o, true); o += 4;
        fr.sp = { n: ns, pos: slice(o, ns * 3, Uint16Array), alive: ns }; o += ns * 6 + pad(ns * 6);
        fr.sp.pk = slice(o, ns, Uint32Array); o += ns * 4;
      }
      if (fl & 4) {
        const nd = dv.getUint32(o, true); o += 4;
        fr.dp = { n: nd, pos: slice(o, nd * 3, Uint16Array) }; o += nd * 6;
        fr.dp.spd = slice(o, nd, Uint8Array); o += nd + pad(nd * 7);
      }
      frames.push(fr);
      maxN = Math.max(maxN, n);
    }
    return Object.assign({}, meta, { frames, maxN: Math.max(maxN, meta.maxN || 0) });
  }
  // version 1 (the native baker): d, fps, frames, objects, ids; per frame n,
  // 4 f32 per object (x y z yaw), 3 u16 + u8 per point
  let o = 8;
  const d = dv.getFloat32(o, true); o += 4;
  const fps = dv.getUint32(o, true); o += 4;
  const nf = dv.getUint32(o, true); o += 4;
  const nobj = dv.getUint32(o, true); o += 4;
  const ids = [];
  for (let k = 0; k < nobj; k++) { ids.push(dv.getUint32(o, true)); o += 4; }
  for (let f = 0; f < nf && o + 4 <= len; f++) {
    const n = dv.getUint32(o, true); o += 4;
    if (o + nobj * 16 + n * 7 > len) break;
    const objs = new Float32Array(nobj * 6);   // the viewer's x y z rx yaw rz; the file's yaw is the other way round
    for (let k = 0; k < nobj; k++) {
      objs[k * 6] = dv.getFloat32(o + k * 16, true); objs[k * 6 + 1] = dv.getFloat32(o + k * 16 + 4, true);
      objs[k * 6 + 2] = dv.getFloat32(o + k * 16 + 8, true); objs[k * 6 + 4] = -dv.getFloat32(o + k * 16 + 12, true);
    }
    o += nobj * 16;
    const pos = slice(o, n * 3, Uint16Array); o += n * 6;
    const spd = slice(o, n, Uint8Array); o += n;
    frames.push({ n, pos, spd, objs });
    maxN = Math.max(maxN, n);
  }
  return { fps, d, frames, maxN, ids, seconds: 0 };
}

module.exports = { packFrame, packDeepFrame, packSprayFrame, packSprayCode, writeCache, readCache };

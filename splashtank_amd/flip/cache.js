'use strict';
// The bake cache on disk (SPLC), so frames baked on the MI355X play in the
// reference's page: packFrame turns a readback into a cache frame (the
// reference's packFrameBuf, src/gpu.js:525-560), writeCache lays frames out as
// version 2 (writeCacheParts, src/shared.js:686-709) and readCache reads
// version 2 and the native baker's version 1 (parseCache2 / parseCache,
// shared.js:710-762; native/splash.cpp:343-360).
//
// Version 2 layout (little endian): 'SPLC', u32 2, u32 JSON length, the JSON
// (spacing, fps, object ids, box, scene, settings), zero padding to 4 bytes,
// u32 frame count; per frame u32 points, u32 flags (1 white water levels
// follow, 2 spray follows, 4 deep water points follow), 6 f32 per object (x y z
// rx yaw rz), 3 u16 per point (position across the box, 0..65535), u8 per
// point (speed in 16ths of m/s times the scene scale, 255 = no point), [u8 per
// point white water level * 17], zero padding to 4 bytes.  This solver makes no
// spray or deep-water points, so flags 2 and 4 are never written (they are read).

const q16 = (v, lo, k) => Math.max(0, Math.min(65535, Math.round((v - lo) * k)));

// readback (id order: x, y, z, w = speed + 64 * foam level) -> cache frame
function packFrame(buf, n, b, states, vs) {
  const ks = 16 * (vs || 1), U = new Uint32Array(buf.buffer, buf.byteOffset, buf.length);
  // water: alive, w below 4096 (deep-water points above) and its sign bit clear (spray)
  const water = (i) => buf[i * 4 + 1] > -1e5 && buf[i * 4 + 3] < 4096 && U[i * 4 + 3] < 0x80000000;
  let last = -1;
  for (let i = 0; i < n; i++) if (water(i)) last = i;
  const m = last + 1;   // empty slots after the last water point are dropped
  const pos = new Uint16Array(m * 3), spd = new Uint8Array(m), fm = new Uint8Array(m);
  const kx = 65535 / (b.x1 - b.x0), ky = 65535 / (b.y1 - b.y0), kz = 65535 / (b.z1 - b.z0);
  let foam = false;
  for (let i = 0; i < m; i++) {
    const o = i * 4, w = buf[o + 3];
    if (!water(i)) { spd[i] = 255; continue; }
    pos[i * 3] = q16(buf[o], b.x0, kx);
    pos[i * 3 + 1] = q16(buf[o + 1], b.y0, ky);
    pos[i * 3 + 2] = q16(buf[o + 2], b.z0, kz);
    const lv = Math.floor(w / 64), s = (w - lv * 64) * ks;
    spd[i] = s > 191 ? 191 : s;
    if (lv) { fm[i] = lv * 17; foam = true; }
  }
  const objs = new Float32Array((states || []).length * 6);
  (states || []).forEach((s, k) => objs.set([s.x, s.y, s.z, s.rx || 0, s.yaw || 0, s.rz || 0], k * 6));
  const fr = { n: m, pos, spd, objs };
  if (foam) fr.fm = fm;
  return fr;
}

// meta: { d, fps, ids, b, org, maxN, seconds, engine, scene, desc, diag, stableIds, scale, spdMs }
function writeCache(meta, frames) {
  const m = { d: meta.d, fps: meta.fps, ids: meta.ids || [], b: meta.b, org: meta.org, maxN: meta.maxN, seconds: meta.seconds,
    engine: meta.engine, scene: meta.scene, desc: meta.desc, diag: meta.diag, stableIds: !!meta.stableIds, scale: meta.scale || 1,
    spdMs: !!meta.spdMs };
  const js = Buffer.from(JSON.stringify(m), 'utf8');
  const pad = (k) => (4 - (k % 4)) % 4;
  const u32 = (...v) => Buffer.from(new Uint32Array(v).buffer);
  const raw = (a) => Buffer.from(a.buffer, a.byteOffset, a.byteLength);
  const parts = [Buffer.from('SPLC', 'latin1'), u32(2, js.length), js, Buffer.alloc(pad(js.length)), u32(frames.length)];
  for (const fr of frames) {
    parts.push(u32(fr.n, fr.fm ? 1 : 0), raw(fr.objs), raw(fr.pos), raw(fr.spd));
    let bytes = fr.n * 7;
    if (fr.fm) { parts.push(raw(fr.fm)); bytes += fr.n; }
    parts.push(Buffer.alloc(pad(bytes)));
  }
  return Buffer.concat(parts);
}

function readCache(buf) {
  const dv = new DataView(buf.buffer, buf.byteOffset, buf.byteLength), len = buf.byteLength;
  if (len < 24 || buf.toString('latin1', 0, 4) !== 'SPLC') throw new Error('not an SPLC cache');
  const slice = (o, n, T) => new T(buf.buffer.slice(buf.byteOffset + o, buf.byteOffset + o + n * T.BYTES_PER_ELEMENT));
  const frames = [];
  let maxN = 0;
  if (dv.getUint32(4, true) === 2) {
    let o = 8;
    const ml = dv.getUint32(o, true); o += 4;
    const meta = JSON.parse(buf.toString('utf8', o, o + ml)); o += ml + ((4 - (ml % 4)) % 4);
    const nf = dv.getUint32(o, true); o += 4;
    const nobj = (meta.ids || []).length;
    for (let f = 0; f < nf && o + 8 <= len; f++) {
      const n = dv.getUint32(o, true), fl = dv.getUint32(o + 4, true); o += 8;
      if (o + nobj * 24 + n * ((fl & 1) ? 8 : 7) > len) break;
      const fr = { n, objs: slice(o, nobj * 6, Float32Array) }; o += nobj * 24;
      fr.pos = slice(o, n * 3, Uint16Array); o += n * 6;
      fr.spd = slice(o, n, Uint8Array); o += n;
      if (fl & 1) { fr.fm = slice(o, n, Uint8Array); o += n; }
      o += (4 - ((n * ((fl & 1) ? 8 : 7)) % 4)) % 4;
      if (fl & 2) {   // spray: u32 count, 3 u16 each, padding, u32 packed each
        const ns = dv.getUint32(o, true); o += 4;
        fr.sp = { n: ns, pos: slice(o, ns * 3, Uint16Array) }; o += ns * 6 + ((4 - ((ns * 6) % 4)) % 4);
        fr.sp.pk = slice(o, ns, Uint32Array); o += ns * 4;
      }
      if (fl & 4) {   // deep water: u32 count, 3 u16 + u8 each, padding
        const nd = dv.getUint32(o, true); o += 4;
        fr.dp = { n: nd, pos: slice(o, nd * 3, Uint16Array) }; o += nd * 6;
        fr.dp.spd = slice(o, nd, Uint8Array); o += nd + ((4 - ((nd * 7) % 4)) % 4);
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

module.exports = { packFrame, writeCache, readCache };

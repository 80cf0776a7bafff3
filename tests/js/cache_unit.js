'use strict';
// The SPLC cache (cache.js) on hand-made frames: byte layout of version 2 at
// fixed offsets, the round trip, and a version 1 file built by hand.  Prints one
// JSON line of results.
const path = require('path');
const C = require(path.join(__dirname, '..', '..', 'splashtank_amd', 'flip', 'cache.js'));

const b = { x0: -0.5, x1: 0.5, y0: 0, y1: 1, z0: -0.5, z1: 0.5 };
// four slots: water (speed 2.5, foam level 3), water (speed 70 -> capped), dead, dead (dropped)
const buf = new Float32Array([
  0, 0.5, 0.25, 2.5 + 64 * 3,
  -0.5, 1, 0.5, 63.5,
  0, -1e6, 0, 0,
  0, -1e6, 0, 0]);
const states = [{ x: 1, y: 2, z: 3, yaw: 0.5 }];
const fr = C.packFrame(buf, 4, b, states, 1);
const out = { n: fr.n, pos: Array.from(fr.pos), spd: Array.from(fr.spd), fm: fr.fm ? Array.from(fr.fm) : null, objs: Array.from(fr.objs) };
const meta = { d: 0.01, fps: 60, ids: [7], b, maxN: 4, seconds: 1.5, engine: 'mi355x' };
const file = C.writeCache(meta, [fr, fr]);
const js = Buffer.from(JSON.stringify({ d: 0.01, fps: 60, ids: [7], b, maxN: 4, seconds: 1.5, engine: 'mi355x', stableIds: false, scale: 1, spdMs: false }), 'utf8');
const dv = new DataView(file.buffer, file.byteOffset, file.byteLength);
const o0 = 12 + js.length + ((4 - (js.length % 4)) % 4);
out.magic = file.toString('latin1', 0, 4);
out.version = dv.getUint32(4, true);
out.jsonLen = [dv.getUint32(8, true), js.length];
out.frames = dv.getUint32(o0, true);
out.frame0 = { n: dv.getUint32(o0 + 4, true), flags: dv.getUint32(o0 + 8, true), yaw: dv.getFloat32(o0 + 12 + 16, true),
  pos0: dv.getUint16(o0 + 36, true), spd1: dv.getUint8(o0 + 36 + 12 + 1), fm0: dv.getUint8(o0 + 36 + 12 + 2) };
const frameBytes = 8 + 24 + 2 * 8 + ((4 - ((2 * 8) % 4)) % 4);
out.length = [file.length, o0 + 4 + 2 * frameBytes];
const back = C.readCache(file);
out.round = back.frames.length === 2 && Array.from(back.frames[1].pos).join() === Array.from(fr.pos).join() &&
  Array.from(back.frames[1].fm).join() === Array.from(fr.fm).join() && back.ids[0] === 7 && back.engine === 'mi355x';
// version 1: 'SPLC', 1, d, fps, frames, objects, ids; frame: n, x y z yaw, pos, spd
const v1 = Buffer.alloc(4 + 4 * 5 + 4 + (4 + 16 + 2 * 7));
v1.write('SPLC', 0, 'latin1'); let o = 4;
v1.writeUInt32LE(1, o); o += 4; v1.writeFloatLE(0.02, o); o += 4; v1.writeUInt32LE(30, o); o += 4; v1.writeUInt32LE(1, o); o += 4;
v1.writeUInt32LE(1, o); o += 4; v1.writeUInt32LE(9, o); o += 4;
v1.writeUInt32LE(2, o); o += 4; [0.1, 0.2, 0.3, 0.75].forEach((x) => { v1.writeFloatLE(x, o); o += 4; });
[1, 2, 3, 4, 5, 6].forEach((x) => { v1.writeUInt16LE(x, o); o += 2; }); v1.writeUInt8(33, o); v1.writeUInt8(44, o + 1);
const r1 = C.readCache(v1);
out.v1 = { fps: r1.fps, n: r1.frames[0].n, yaw: r1.frames[0].objs[4], pos: Array.from(r1.frames[0].pos), spd: Array.from(r1.frames[0].spd), ids: r1.ids };
console.log(JSON.stringify(out));

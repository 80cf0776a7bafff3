'use strict';
// Starting water for the facade: the narrow-band start nbSeed (src/shared.js:805-853),
// water volumes from a mesh's distance grid (fillVolume, shared.js:208-218, with
// sdfSample, shared.js:183-205), and the ocean's surface and velocity in f64 for
// seeding (oceanEtaJS / oceanVelJS, shared.js:792-803).
//
// nbSeed puts particles only in a band under each shape's free surface (box
// faces on the tank walls, and water-mesh surface points on a wall, are not
// free surfaces) and under the ocean's waves outside the outer half of its
// boundary layer; the deep water below starts on the solver's grid from the same
// shapes (FlipSim's nbFills / nbFillVols, st_params.fills).  Spare parked slots
// follow, for the solver's reseed to refill the band from.  Jitter comes from
// sim.rand (Math.random unless the caller seeds it), called x, y, z per particle
// as the reference's Math.random calls are.

// Trilinear distance of a grid {dims, origin, cell, data} at a local point, with
// the distance to the grid added outside it; out = [phi, nx, ny, nz] (unit normal)
function sdfSample(sd, x, y, z, out) {
  const [nx, ny, nz] = sd.dims, c = sd.cell, inv = 1 / c;
  const g = [(x - sd.origin[0]) * inv, (y - sd.origin[1]) * inv, (z - sd.origin[2]) * inv];
  const hi = [nx - 1.0001, ny - 1.0001, nz - 1.0001], off = [0, 0, 0];
  for (let a = 0; a < 3; a++) {
    if (g[a] < 0) { off[a] = g[a]; g[a] = 0; } else if (g[a] > hi[a]) { off[a] = g[a] - hi[a]; g[a] = hi[a]; }
  }
  const i = g[0] | 0, j = g[1] | 0, k = g[2] | 0, fx = g[0] - i, fy = g[1] - j, fz = g[2] - k;
  const D = sd.data, base = i + nx * (j + ny * k), sl = nx * ny;
  const v = (di, dj, dk) => D[base + di + nx * dj + sl * dk];
  const lerp = (a, b2, t) => a + (b2 - a) * t;
  const c00 = lerp(v(0, 0, 0), v(1, 0, 0), fx), c10 = lerp(v(0, 1, 0), v(1, 1, 0), fx);
  const c01 = lerp(v(0, 0, 1), v(1, 0, 1), fx), c11 = lerp(v(0, 1, 1), v(1, 1, 1), fx);
  const c0 = lerp(c00, c10, fy), c1 = lerp(c01, c11, fy);
  const phi = lerp(c0, c1, fz);
  const e2 = off[0] * off[0] + off[1] * off[1] + off[2] * off[2];
  if (e2 > 0) {   // outside the grid: the distance to it along the way out
    const l = Math.sqrt(e2);
    out[0] = phi + l * c; out[1] = off[0] / l; out[2] = off[1] / l; out[3] = off[2] / l;
    return out;
  }
  const gx = ((v(1, 0, 0) - v(0, 0, 0)) * (1 - fy) + (v(1, 1, 0) - v(0, 1, 0)) * fy) * (1 - fz) +
    ((v(1, 0, 1) - v(0, 0, 1)) * (1 - fy) + (v(1, 1, 1) - v(0, 1, 1)) * fy) * fz;
  const gy = (c10 - c00) * (1 - fz) + (c11 - c01) * fz, gz = c1 - c0;
  const l = Math.hypot(gx, gy, gz) || 1;
  out[0] = phi; out[1] = gx / l; out[2] = gy / l; out[3] = gz / l;
  return out;
}

// a water mesh at its pose: the local distance of a world point (R row major, local -> world)
function localDistance(f, x, y, z, out) {
  const R = f.R, dx = x - f.x, dy = y - f.y, dz = z - f.z;
  return sdfSample(f.sdf, R[0] * dx + R[3] * dy + R[6] * dz, R[1] * dx + R[4] * dy + R[7] * dz, R[2] * dx + R[5] * dy + R[8] * dz, out);
}

// every lattice point inside the posed mesh gets a particle at rest (y outer, x, z inner)
function fillVolume(sim, f) {
  const d = sim.d, jit = 0.01 * d / 0.12, rnd = sim.rand || Math.random, tmp = [0, 0, 0, 0];
  for (let y = f.lo[1] + d / 2; y < f.hi[1]; y += d)
    for (let x = f.lo[0] + d / 2; x < f.hi[0]; x += d)
      for (let z = f.lo[2] + d / 2; z < f.hi[2]; z += d) {
        if (localDistance(f, x, y, z, tmp)[0] >= 0) continue;
        const px = x + (rnd() - 0.5) * jit, py = y + (rnd() - 0.5) * jit, pz = z + (rnd() - 0.5) * jit;
        if (!sim.add(px, py, pz, 0, 0, 0)) return;
      }
}

// the ocean's wave list W (Float32Array, 8 per wave) at a point in f64: surface
// height level + sum a cos(theta), and the Airy velocity under it over a floor at
// y0, the height stretched so the crest takes the surface value
function oceanEta(W, level, x, z, t) {
  let e = level;
  for (let k = 0; k < W.length / 8; k++) {
    const w = W.subarray(k * 8, k * 8 + 8);
    e += w[2] * Math.cos(w[0] * x + w[1] * z - w[3] * t + w[4]);
  }
  return e;
}
function oceanVel(W, level, y0, x, y, z, t, out) {
  const D = Math.max(level - y0, 1e-3), eta = oceanEta(W, level, x, z, t);
  const ys = Math.min(Math.max((y - y0) * D / Math.max(D + eta - level, 0.1 * D), 0), D);
  let u = 0, v = 0, s = 0;
  for (let k = 0; k < W.length / 8; k++) {
    const w = W.subarray(k * 8, k * 8 + 8), kk = Math.max(w[5], 1e-4);
    const th = w[0] * x + w[1] * z - w[3] * t + w[4];
    const e1 = Math.exp(kk * (ys - D)), e2 = Math.exp(-kk * (ys + D)), dn = Math.max(1 - Math.exp(-2 * kk * D), 1e-4);
    const aw = w[2] * w[3], hc = aw * (e1 + e2) / dn * Math.cos(th) / kk;
    u += hc * w[0]; v += aw * (e1 - e2) / dn * Math.sin(th); s += hc * w[1];
  }
  out[0] = u; out[1] = v; out[2] = s;
  return out;
}

// nbSeed(sim, fills, nb, oc, t, g, mode): mode 1 a restart, 2 water added to a
// running sim.  fills: {t:'box', x0..z1} | {t:'sphere', x, y, z, r, vy} | {t:'vol',
// sdf, x, y, z, R, lo, hi}.  Returns {seeded, full}.
function nbSeed(sim, fills, nb, oc, t, g, mode) {
  const d = sim.d, h = sim.h, b = sim.b, band = Math.max(2, (nb && nb.band) || 3) * h;
  const jit = 0.01 * d / 0.12, near = 0.6 * h, n0 = sim.N, rnd = sim.rand || Math.random;
  const prims = [], vols = [];
  let full = false;
  const put = (x, y, z, vx, vy, vz) => {
    const px = x + (rnd() - 0.5) * jit, py = y + (rnd() - 0.5) * jit, pz = z + (rnd() - 0.5) * jit;
    if (!sim.add(px, py, pz, vx, vy, vz)) full = true;
  };
  const onWall = (x, y, z) => x < b.x0 + near || x > b.x1 - near || z < b.z0 + near || z > b.z1 - near ||
    y < b.y0 + near || y > b.y1 - near;
  // a face within `near` of its wall is no free surface: pushed far out
  const away = (v, wall, dirOut) => ((dirOut < 0 ? v <= wall + near : v >= wall - near) ? dirOut * 1e4 : v);
  for (const f of fills) {
    if (full) break;
    if (f.t === 'box') {
      const e = { x0: away(f.x0, b.x0, -1), x1: away(f.x1, b.x1, 1), y0: away(f.y0, b.y0, -1),
        y1: away(f.y1, b.y1, 1), z0: away(f.z0, b.z0, -1), z1: away(f.z1, b.z1, 1) };
      for (let y = f.y0 + d / 2; y < f.y1 && !full; y += d)
        for (let x = f.x0 + d / 2; x < f.x1 && !full; x += d)
          for (let z = f.z0 + d / 2; z < f.z1; z += d)
            if (Math.min(x - e.x0, e.x1 - x, y - e.y0, e.y1 - y, z - e.z0, e.z1 - z) <= band) put(x, y, z, 0, 0, 0);
      prims.push(Object.assign({ t: 'box' }, e));
    } else if (f.t === 'vol') {
      const tmp = [0, 0, 0, 0], R = f.R;
      const inside = (x, y, z) => localDistance(f, x, y, z, tmp)[0] < 0;
      for (let y = f.lo[1] + d / 2; y < f.hi[1] && !full; y += d)
        for (let x = f.lo[0] + d / 2; x < f.hi[0] && !full; x += d)
          for (let z = f.lo[2] + d / 2; z < f.hi[2]; z += d) {
            const s = localDistance(f, x, y, z, tmp)[0];
            if (s >= 0 || -s > band) continue;
            // the nearest surface point in world space; on a wall it is no free surface, so the depth is measured upwards
            const wn = [R[0] * tmp[1] + R[1] * tmp[2] + R[2] * tmp[3], R[3] * tmp[1] + R[4] * tmp[2] + R[5] * tmp[3],
              R[6] * tmp[1] + R[7] * tmp[2] + R[8] * tmp[3]];
            if (!onWall(x - s * wn[0], y - s * wn[1], z - s * wn[2])) { put(x, y, z, 0, 0, 0); continue; }
            let up = 0;
            while (up <= band && y + up < b.y1 && inside(x, y + up, z)) up += d;
            if (up <= band && y + up < b.y1) put(x, y, z, 0, 0, 0);
          }
      vols.push({ sdf: f.sdf, x: f.x, y: f.y, z: f.z, R: f.R, vx: 0, vy: 0, vz: 0, wx: 0, wy: 0, wz: 0 });
    } else {
      const r = f.r;
      for (let y = -r; y <= r && !full; y += d)
        for (let x = -r; x <= r && !full; x += d)
          for (let z = -r; z <= r; z += d) {
            const depth = r - Math.hypot(x, y, z);
            if (depth >= 0 && depth <= band) put(f.x + x, f.y + y, f.z + z, 0, f.vy || 0, 0);
          }
      prims.push({ t: 'sphere', x: f.x, y: f.y, z: f.z, r });
    }
  }
  if (oc && !full) {   // the band under the waves, moving with them, outside the outer half of the boundary
    const W = sim.oceanWaves(oc, g), v = [0, 0, 0], lay = Math.max(oc.layer, 1e-4);
    for (let x = b.x0 + d / 2; x < b.x1 && !full; x += d)
      for (let z = b.z0 + d / 2; z < b.z1 && !full; z += d) {
        if (1 - Math.min(x - b.x0, b.x1 - x, z - b.z0, b.z1 - z) / lay >= 0.5) continue;
        const eta = Math.min(oceanEta(W, oc.level, x, z, t), b.y1 - d);
        const y1 = b.y0 + d / 2 + Math.max(0, Math.floor((eta - band - b.y0) / d)) * d;
        for (let y = y1; y < eta; y += d) { oceanVel(W, oc.level, b.y0, x, y, z, t, v); put(x, y, z, v[0], v[1], v[2]); }
      }
  }
  const seeded = sim.N - n0, spare = Math.min(sim.cap - sim.N, Math.ceil(0.35 * seeded) + 4000);
  for (let k = 0; k < spare && sim.N < sim.cap; k++) sim.add(0, -1e6, 0, 0, 0, 0);
  if (mode === 2) { sim.nbFills = (sim.nbFills || []).concat(prims); sim.nbFillVols = (sim.nbFillVols || []).concat(vols); }
  else { sim.nbFills = prims; sim.nbFillVols = vols; }
  sim.nbInit = mode || 1; sim.nbCfg = nb; sim.oceanCfg = oc;
  return { seeded, full };
}

module.exports = { nbSeed, fillVolume, sdfSample, oceanEta, oceanVel };

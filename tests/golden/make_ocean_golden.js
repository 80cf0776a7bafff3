'use strict';
// Writes the golden vectors of the reference's ocean: the wave lists of
// oceanWaves (src/shared.js:769-791) for several ocean settings, and the surface
// height and water velocity of its CPU twins oceanEtaJS / oceanVelJS
// (src/shared.js:792-803, f64) at sample points and times.  The settings are
// the ones of the reference's own "Meteor lake" preset (src/presets.js) and
// unit-box variants covering the seed rule, shallow water, a calm sea and the
// amplitude cap.  Run once in a container that holds the reference:
//   node tests/golden/make_ocean_golden.js /root/reference/src/shared.js tests/golden
// Output: ocean_golden.json.
const fs = require('fs');
const vm = require('vm');
const path = require('path');
const [src, outDir] = process.argv.slice(2);
const ctx = vm.createContext({ Math, Float32Array, Float64Array, Int32Array, Uint32Array, Array, Map, Object, JSON, Number, console });
vm.runInContext(fs.readFileSync(src, 'utf8') + '\n;this.__api = { oceanWaves, oceanEtaJS, oceanVelJS };', ctx);
const api = ctx.__api;
const cases = [
  { name: 'meteor_lake', oc: { on: true, height: 0.15, length: 30, dir: 0, spread: 25, level: 0.1, layer: 0.06, seed: 1 }, g: 9.8, y0: 0 },
  { name: 'unit_box', oc: { on: true, height: 0.04, length: 0.45, dir: 30, spread: 40, level: 0.35, layer: 0.12, seed: 7 }, g: 9.81, y0: 0 },
  { name: 'seed_zero', oc: { on: true, height: 0.04, length: 0.45, dir: -20, spread: 10, level: 0.35, layer: 0.12, seed: 0 }, g: 9.81, y0: 0 },
  { name: 'seed_negative', oc: { on: true, height: 0.03, length: 0.3, dir: 90, spread: 0, level: 0.3, layer: 0.1, seed: -123456 }, g: 9.81, y0: 0 },
  { name: 'shallow', oc: { on: true, height: 0.02, length: 1.5, dir: 10, spread: 20, level: 0.05, layer: 0.1, seed: 3 }, g: 9.81, y0: 0 },
  { name: 'calm', oc: { on: true, height: 0, length: 0.5, dir: 0, spread: 30, level: 0.3, layer: 0.1, seed: 5 }, g: 9.81, y0: 0 },
  { name: 'capped', oc: { on: true, height: 0.6, length: 0.4, dir: 45, spread: 60, level: 0.4, layer: 0.1, seed: 11 }, g: 9.81, y0: 0 },
  { name: 'big_seed', oc: { on: true, height: 0.05, length: 0.6, dir: 200, spread: 15, level: 0.3, layer: 0.1, seed: 2000000000 }, g: 9.81, y0: 0 },
];
const out = [];
for (const c of cases) {
  const depth = c.oc.level - c.y0;
  const W = api.oceanWaves(c.oc, c.g, depth);
  const samples = [];
  for (const t of [0, 0.37, 4.5]) {
    for (const [x, z] of [[-0.41, 0.13], [0.2, -0.33], [0.47, 0.46]]) {
      const eta = api.oceanEtaJS(W, c.oc, x, z, t);
      for (const fy of [0.1, 0.6, 1.0]) {
        const y = c.y0 + fy * depth;
        const v = api.oceanVelJS(W, c.oc, c.y0, x, y, z, t, [0, 0, 0]);
        samples.push({ x, y, z, t, eta, v: v.slice() });
      }
    }
  }
  out.push({ name: c.name, oc: c.oc, g: c.g, depth, n: W.n, waves: Array.from(W.w), samples });
}
fs.writeFileSync(path.join(outDir, 'ocean_golden.json'),
  JSON.stringify({ source: 'oceanWaves, oceanEtaJS, oceanVelJS (src/shared.js:769-803)', cases: out }));
console.log(out.map((c) => `${c.name}: ${c.n} waves, ${c.samples.length} samples`).join('\n'));

'use strict';
// The MI355X FlipSim for Node: the simulator (sim.js) and its scene driving
// (runner.js: SceneRunner, keyframed poses, the adaptive-substep bake loop) and
// the SPLC bake cache (cache.js), and the collider distance grids (sdf.js).
module.exports = Object.assign({}, require('./sim.js'), require('./runner.js'), require('./cache.js'), require('./sdf.js'));

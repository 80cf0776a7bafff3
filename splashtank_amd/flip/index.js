'use strict';
// The MI355X FlipSim for Node: see sim.js.
module.exports = require('./sim.js');

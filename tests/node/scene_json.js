'use strict';
// A scene, its parameters and grips written by tests/node_addon.py as JSON, turned into the typed
// arrays the addon reads (the CompiledScene layout of the reference, scene.ts:91-125).
const fs = require('fs');

const TYPES = {
  position: Float64Array, mass: Float64Array, pinned: Uint8Array, restLength: Float64Array, softening: Float64Array,
  segFlags: Uint16Array, intact: Float64Array, cutMask: Uint32Array, failSite: Float64Array,
  crossingNormal0: Float64Array, tails: Float64Array, fieldStrength: Float32Array, fieldDamage: Uint8Array,
  fieldArea: Uint8Array,
};

function load(file) {
  const j = JSON.parse(fs.readFileSync(file, 'utf8'));
  const scene = { nodeCount: j.scene.nodeCount, crossingCount: j.scene.crossingCount, layers: j.scene.layers };
  for (const [k, T] of Object.entries(TYPES)) scene[k] = T.from(j.scene[k]);
  const grips = j.grips ? {
    node: Int32Array.from(j.grips.node), grip: Int32Array.from(j.grips.grip), weight: Float64Array.from(j.grips.weight),
    offset: Float64Array.from(j.grips.offset), goal: Float64Array.from(j.grips.goal),
    stiffness: Float64Array.from(j.grips.stiffness),
  } : null;
  return { scene, params: j.params, grips, goals: j.goals || [], steps: j.steps || 1 };
}

module.exports = { load };

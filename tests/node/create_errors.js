'use strict';
// The constructor's error paths through the N-API addon, without a GPU: the library validates the
// scene and parameters before it looks for a device, and every failure is a JS Error with the
// library's message (the reference's GpuSolver throws for its limits, gpu/solver.ts:298-363).
const assert = require('assert');
const path = require('path');
const { GpuSolver, BD_MAX_LINE } = require(path.join(__dirname, '..', '..', 'break-that-denim_amd', 'node', 'gpu_solver.js'));

function yarn() {
  return { linearDensity: 1e-4, eaLow: 20, eaHigh: 20, crimpStrain: 0.01, eaCompress: 20, axialViscosity: 1e-4,
    bendRigidity: 1e-8, bendViscosity: 1e-3, strength: 1e9, softeningOpening: 1e-4, yarnsPerSpan: 1 };
}

function sheet(nx, ny) {
  const N = nx * ny;
  const material = { pitchU: 1e-3, pitchV: 1e-3, yarns: [yarn(), yarn()], shearStiffness: 8e3, crossingFriction: 0.3,
    crossingNormal0: 0.1, crossingWrap: 0.2, airDrag: 1.5, thickness: 7e-4 };
  const scene = { nodeCount: N, crossingCount: N, layers: [{ nx, ny, nodeOffset: 0, bundleU: 1, bundleV: 1, fieldOffset: 0,
    fieldK: nx, fieldM: ny, weavePitchU: 1e-3, weavePitchV: 1e-3, material }] };
  const arrays = { position: [Float64Array, 3], mass: [Float64Array, 1], pinned: [Uint8Array, 1], restLength: [Float64Array, 2],
    softening: [Float64Array, 2], segFlags: [Uint16Array, 2], intact: [Float64Array, 2], cutMask: [Uint32Array, 2],
    failSite: [Float64Array, 2], crossingNormal0: [Float64Array, 1], tails: [Float64Array, 4], fieldStrength: [Float32Array, 2],
    fieldDamage: [Uint8Array, 2], fieldArea: [Uint8Array, 2] };
  for (const [k, [T, per]] of Object.entries(arrays)) scene[k] = new T(per * N);
  for (let n = 0; n < N; n++) {
    scene.position[3 * n] = (n % nx) * 1e-3;
    scene.position[3 * n + 2] = Math.floor(n / nx) * 1e-3;
    scene.mass[n] = 3.5e-7;
  }
  return scene;
}

function throws(fn, pattern, label) {
  let err = null;
  try { fn(); } catch (e) { err = e; }
  assert(err instanceof Error, `${label}: no error thrown`);
  assert(pattern.test(err.message), `${label}: unexpected message "${err.message}"`);
  console.log(`ok ${label}: ${err.message}`);
}

throws(() => new GpuSolver(sheet(BD_MAX_LINE + 1, 4), {}), /BD_MAX_LINE/, 'line beyond BD_MAX_LINE');
throws(() => new GpuSolver(sheet(8, 8), { newtonIterations: 0 }), /newton_iterations/, 'zero Newton iterations');
const bad = sheet(8, 8);
bad.nodeCount = 63;
throws(() => new GpuSolver(bad, {}), /node_count|typed array/, 'inconsistent node count');
const missing = sheet(8, 8);
delete missing.position;
throws(() => new GpuSolver(missing, {}), /scene\.position/, 'missing position array');
// a valid scene: created on a GPU host, refused for want of a device on a CPU host
let solver = null;
try {
  solver = new GpuSolver(sheet(8, 8), {});
  console.log('ok valid scene: created on the GPU');
} catch (e) {
  assert(/no HIP device/.test(e.message), `valid scene: unexpected message "${e.message}"`);
  console.log(`ok valid scene: ${e.message}`);
}
console.log('create_errors: all checks passed');

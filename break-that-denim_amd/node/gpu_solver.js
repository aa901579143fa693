'use strict';
/*
 * GpuSolver over the N-API addon (build/Release/bd.node, built by node-gyp from bd_node.c): the
 * Node host's drop-in for the reference's WebGPU GpuSolver (src/simulation/gpu/solver.ts:194-838).
 * Same member names and argument meaning; every call runs the MI355X HIP kernels of libbd.so.
 *
 *   const { GpuSolver } = require('./gpu_solver');
 *   const solver = new GpuSolver(scene, params);   // scene: CompiledScene-shaped typed arrays
 *   solver.setState(x4, v4);                       // vec4 per node, as the reference's setState
 *   solver.encodeSubstep();                        // one substep, enqueued on the solver's stream
 *   const x = solver.readPositions();              // Float64Array(3N)
 *
 * Errors are thrown as JS Errors carrying the library's message, as the reference throws for its
 * limits (gpu/solver.ts:298-363).
 */
const path = require('path');

const addon = require(path.join(__dirname, 'build', 'Release', 'bd.node'));

const BUF = { X: 0, V: 1, DX: 2, PRED: 3, DELTA: 4, RHS: 5, REST: 12, INTACT: 13, CUTMASK: 14, SEGFLAGS: 15,
  FAILSITE: 16, TAILS: 17, FIELD: 18 };

// vec4 per node (the reference's layout, w lane unused here) -> xyz doubles
function xyz(v4, N) {
  const out = new Float64Array(3 * N);
  const stride = v4.length >= 4 * N ? 4 : 3;
  for (let n = 0; n < N; n++) for (let q = 0; q < 3; q++) out[3 * n + q] = v4[stride * n + q];
  return out;
}

class GpuSolver {
  /** new GpuSolver(device, scene, params) (gpu/solver.ts:275); the device is the process's GPU. */
  constructor(scene, params = {}) {
    this.nodeCount = scene.nodeCount;
    this.crossingCount = scene.crossingCount;
    this.handle = addon.create(scene, params);
  }

  /** encodeSubstep(encoder, minIterations, maxIterations) (gpu/solver.ts:737-760). */
  encodeSubstep(minIterations = 0, maxIterations = 0) { addon.substep(this.handle, minIterations, maxIterations); }

  /** setState(x4, v4?) (gpu/solver.ts:831-836): vec4 (or xyz) per node. */
  setState(x4, v4) {
    addon.setState(this.handle, xyz(x4, this.nodeCount), v4 ? xyz(v4, this.nodeCount) : null);
  }

  /** The ReferenceGrip springs (solver.ts:25-37): arrays per entry and per grip. */
  setGrips(grips) { addon.setGrips(this.handle, grips); }

  /** The grips' goals for the following substeps (3 per grip). */
  setGripGoals(goals) { addon.setGripGoals(this.handle, Float64Array.from(goals)); }

  /** Newton solves with the learned coarse-space field: 0 (or false) plain MG-PCG, 1 (or true) a
   *  warm start from the predicted fields, 2 deflated flexible PCG over them. */
  setLearned(mode) { addon.setLearned(this.handle, mode === true ? 1 : mode === false ? 0 : mode | 0); }
  learnRecord(capacity) { addon.learnRecord(this.handle, capacity); }
  learnTrain(epochs, lr = 1e-3, batch = 1, targets = 3) { return addon.learnTrain(this.handle, epochs, lr, batch, targets); }

  readPositions() { const a = new Float64Array(3 * this.nodeCount); addon.read(this.handle, BUF.X, 0, a); return a; }
  readVelocities() { const a = new Float64Array(3 * this.nodeCount); addon.read(this.handle, BUF.V, 0, a); return a; }
  readField() { const a = new Uint8Array(2 * this.crossingCount); addon.read(this.handle, BUF.FIELD, 0, a); return a; }
  stats() { return addon.stats(this.handle); }
  sync() { addon.sync(this.handle); }
}

module.exports = { GpuSolver, BUF, BD_MAX_LINE: addon.BD_MAX_LINE };

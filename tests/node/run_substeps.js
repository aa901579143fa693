'use strict';
// node run_substeps.js in.json out.json: the scene's substeps through the N-API addon (goals per
// step from the JSON), then the positions, velocities and per-step statistics as JSON.
const fs = require('fs');
const path = require('path');
const { GpuSolver } = require(path.join(__dirname, '..', '..', 'break-that-denim_amd', 'node', 'gpu_solver.js'));
const { load } = require('./scene_json');

const [inFile, outFile] = process.argv.slice(2);
const { scene, params, grips, goals, steps } = load(inFile);
const solver = new GpuSolver(scene, params);
if (grips) solver.setGrips(grips);
const stats = [];
for (let k = 0; k < steps; k++) {
  if (goals[k]) solver.setGripGoals(goals[k]);
  solver.encodeSubstep();
  stats.push(solver.stats());
}
fs.writeFileSync(outFile, JSON.stringify({ x: Array.from(solver.readPositions()), v: Array.from(solver.readVelocities()), stats }));
console.log(`node: ${steps} substeps of ${scene.nodeCount} nodes, pcg ${JSON.stringify(stats.map((s) => s.pcgIterations))}`);

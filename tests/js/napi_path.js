'use strict';
// The thermal path kept on the device through the Node-API addon: two steps of
// a plan solved and kept, their sum against one solve over the whole interval
// (temperature-independent materials: the thermal stress is linear in T), the
// single-precision increments read back, and a restore of those increments
// summed on the device against the same sum in JavaScript.
//   node tests/js/napi_path.js [spec name]   -> one JSON line
const fs = require('fs');
const path = require('path');
const fh = require('../../packaging-tm-modeling-tool_amd/js/fcbga_hip.js');
const { buildMesh, groupsFor } = require('./spec_mesh.js');

const GOLDEN = path.join(__dirname, '..', 'golden');
const name = process.argv[2] || 'plate_16';
const spec = JSON.parse(fs.readFileSync(path.join(GOLDEN, name + '_spec.json'), 'utf8'));
const mesh = buildMesh(spec);
const solver = new fh.HipStepSolver(mesh, { tol: 1e-12, precision: 'fp64' });
// the thermal stress of each step is that of its own interval
const step = (Ta, Tb) => fh.stepFromGroups(Ta, Tb, 0, groupsFor(Object.assign({}, spec, { Ta, Tb, stage: 0 })));
solver.pathReset();
const r0 = solver.pathSolve(0, step(217, 150), '217 -> 150');
const r1 = solver.pathSolve(1, step(150, 25), '150 -> 25');
const sum = solver.pathSum([0, 1]);
const one = solver.solveStep(step(217, 25), '217 -> 25').u;
let e1 = 0, mx = 0;
for (let q = 0; q < one.length; q++) { e1 = Math.max(e1, Math.abs(sum[q] - one[q])); mx = Math.max(mx, Math.abs(one[q])); }
// single precision read-back, against the kept FP64 increment
const d0 = solver.pathDu(0), d1 = solver.pathDu(1), p1 = solver.pathSum([1]);
let f32ok = d0.length === one.length && d1.length === one.length;
for (let q = 0; q < p1.length; q++) f32ok = f32ok && d1[q] === Math.fround(p1[q]);
// restore and sum on the device, against the same FP64 sum of the FP32 values here
solver.pathReset();
solver.pathLoad(0, d0);
solver.pathLoad(1, d1);
const back = solver.pathSum([0, 1]);
let exact = true;
for (let q = 0; q < back.length; q++) exact = exact && back[q] === (0 + d0[q]) + d1[q];
// an FP64 increment (a helper's partial message) kept as it is
solver.pathLoad(2, p1);
const p2 = solver.pathSum([2]);
for (let q = 0; q < p1.length; q++) exact = exact && p2[q] === p1[q];
let badLoad = false;
try { solver.pathLoad(3, new Float32Array(3)); } catch (e) { badLoad = /Float64Array of ndof/.test(e.message); }
solver.destroy();
console.log(JSON.stringify({ name, ndof: one.length, converged: r0.converged && r1.converged, rel_err_sum: e1 / mx, f32ok, restore_exact: exact, badLoad }));

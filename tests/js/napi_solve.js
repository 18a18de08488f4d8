'use strict';
// Step solve through the Node-API addon for committed mesh specs, checked
// against the reference's golden displacements (mesh and step data from
// spec_mesh.js).
//   node tests/js/napi_solve.js <spec name> [mixed|fp64] [mg|ic]   -> one JSON line
const fs = require('fs');
const path = require('path');
const zlib = require('zlib');
const fh = require('../../packaging-tm-modeling-tool_amd/js/fcbga_hip.js');

const GOLDEN = path.join(__dirname, '..', 'golden');

const { buildMesh, groupsFor } = require('./spec_mesh.js');

const name = process.argv[2] || 'plate_16';
const precision = process.argv[3] || 'mixed';
const precond = process.argv[4] || 'mg';
const spec = JSON.parse(fs.readFileSync(path.join(GOLDEN, name + '_spec.json'), 'utf8'));
const mesh = buildMesh(spec);
const solver = new fh.HipStepSolver(mesh, { tol: 1e-10, precision, precond });
const rec = solver.solveStep(fh.stepFromGroups(spec.Ta, spec.Tb, spec.stage || 0, groupsFor(spec)), name);
const buf = zlib.gunzipSync(fs.readFileSync(path.join(GOLDEN, name + '_u.f64.gz')));
const uref = new Float64Array(buf.buffer, buf.byteOffset, buf.length / 8);
let err = 0, mx = 0;
for (let q = 0; q < uref.length; q++) { err = Math.max(err, Math.abs(rec.u[q] - uref[q])); mx = Math.max(mx, Math.abs(uref[q])); }
solver.destroy();
console.log(JSON.stringify({ name, precision, ndof: rec.u.length, iters: rec.iters, converged: rec.converged, relres: rec.relres, rel_err: err / mx, ms: rec.ms,
  precond: rec.precond, fallback: rec.fallback, icShift: rec.icShift }));

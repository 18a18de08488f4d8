// Golden vectors from the reference JS core (run in the build container only:
//   node tests/golden/gen_golden.js [/root/reference] [anand|element|stress|warm|ic|screen|submodel|path|<spec name> ...]
// ). It loads src/core/*.js the way the reference's own tools/dev-core.mjs does,
// builds each committed mesh spec (*_spec.json) with the reference mesher's
// buildStructuredMesh, assembles one thermal step and solves it with the
// reference LinearSolver at a tight tolerance, then writes:
//   <name>_u.f64.gz, <name>_rhs.f64.gz  (Float64 little-endian, gzip)
//   <name>_golden.json                  (solve record, warpage, sizes)
//   element_golden.json                 (condensed element tables on sampled boxes)
//   anand_golden.json                   (Anand material-point step sequences)
//   <name>_stress_golden.json           (element states, point stresses, die surface maps, joint loads;
//                                        cfg2 with the argument stress_cfg2, a three-minute reference solve)
//   <name>_warm_golden.json, <name>_warm_u.f64.gz (warm-started solves along a path with T-dependent moduli)
//   <name>_ic_golden.json, <name>_ic_u.f64.gz     (IC(0)-preconditioned solves)
//   fallback_golden.json                (the fallback to the other preconditioner)
//   screen_golden.json, screen_<case>_table.f64.gz (bump screening on synthetic strain tables, cycle histories, site subsampling)
//   submodel_golden.json, submodel_*.f64.gz (a bump submodel of the bare-die preset: set-up data and result)
//   path_golden.json                    (thermal path plans, evaluation temperatures and step sums)
//   package_<bare|lidded>_golden.json.gz (reference-built Draft package meshes with layered sections, two steps)
//   package_modules_golden.json.gz      (argument package_modules: a stacked-module package, module-stage steps;
//   package_modules_path.json.gz         with its whole thermal path, step data and 25 C field)
//   tornado_golden.json.gz, tornado_blocks.f64.gz (arguments tornado [nproc], then tornado_screen: the
//                                        reference's tornado cases with their plans, step data and meshes)
// Only data leaves this script; no reference source is copied.
'use strict';
const fs = require('fs');
const path = require('path');
const zlib = require('zlib');

const REF = process.argv[2] || '/root/reference';
const HERE = __dirname;
const dir = path.join(REF, 'src', 'core');
const src = fs.readdirSync(dir).filter(f => f.endsWith('.js')).sort().map(f => fs.readFileSync(path.join(dir, f), 'utf8')).join('\n');
const get = new Function(src + '\nreturn (name) => eval(name);')();

const buildStructuredMesh = get('buildStructuredMesh'), constraints321 = get('constraints321');
const FEModel = get('FEModel'), LinearSolver = get('LinearSolver');
const shapeMatrices = get('shapeMatrices'), elementK33Fast = get('elementK33Fast'), elementFth = get('elementFth');
const condense33 = get('condense33'), isoD = get('isoD'), columnD = get('columnD');
const warpageJEITA = get('warpageJEITA'), bgaSites = get('bgaSites');

const f64gz = (arr, file) => fs.writeFileSync(path.join(HERE, file), zlib.gzipSync(Buffer.from(Float64Array.from(arr).buffer), { level: 9 }));
const list = a => Array.from(a);

function testMaterial(m) {
  return { id: m.name, name: m.name, elastic: { type: 'iso', E: { type: 'const', E: { v: m.E / 1000, u: 'GPa' } }, nu: { v: m.nu } }, cte: { type: 'const', a: { v: m.cte } } };
}

function assignFor(spec) {
  const xs = spec.xs, ys = spec.ys;
  const inside = (r, x, y) => x >= r[0] && x <= r[1] && y >= r[2] && y <= r[3];
  return (i, j, k) => {
    const xc = (xs[i] + xs[i + 1]) / 2, yc = (ys[j] + ys[j + 1]) / 2;
    for (const r of spec.cells) {
      if (!(r.k0 <= k && k < r.k1)) continue;
      if (r.in && !inside(r.in, xc, yc)) continue;
      if (r.out && inside(r.out, xc, yc)) continue;
      const mats = spec.materials;
      const out = [{ group: r.mat, scale: r.scale == null ? 1 : r.scale, part: mats[r.mat].part }];
      if (r.also != null) out.push({ group: r.also, scale: r.also_scale == null ? 1 : r.also_scale, part: mats[r.also].part });
      return out;
    }
    return [];
  };
}

function solveSpec(name) {
  const spec = JSON.parse(fs.readFileSync(path.join(HERE, name + '_spec.json'), 'utf8'));
  const groups = spec.materials.map(m => ({ name: m.name, kind: 'homog', mat: m.name, stage: m.stage || 0, part: m.part, column: !!m.column }));
  const mesh = buildStructuredMesh(spec.xs, spec.ys, spec.zs, groups, assignFor(spec));
  mesh.constraints = constraints321(mesh, spec.constraint_span);
  if (spec.feature_x && spec.feature_x.length) { mesh.featureX = Int32Array.from(spec.feature_x); mesh.featureY = Int32Array.from(spec.feature_y); }
  const lib = spec.materials.map(testMaterial);
  const model = new FEModel(mesh, { materials: lib });
  const solver = new LinearSolver(model, { precond: 'mg', tol: spec.golden_tol || 1e-12 });
  const rhs = new Float64Array(model.ndof);
  const t0 = Date.now();
  model.assembleStep(spec.Ta, spec.Tb, spec.stage || 0, rhs);
  const rhsCopy = Float64Array.from(rhs);
  const res = solver.solve(rhs, name);
  const ms = Date.now() - t0;
  const cfg = { bga: { pitch: spec.bga.pitch, inset: spec.bga.inset }, substrate: { sx: spec.bga.sx, sy: spec.bga.sy }, limits: {} };
  const geom = { bga: bgaSites(cfg) };
  const w = warpageJEITA(mesh, geom, res.u, cfg);
  // top plane w at every existing node of that plane (i, j order)
  const top = [], kt = spec.top_plane;
  for (let i = 0; i < mesh.nx; i++) for (let j = 0; j < mesh.ny; j++) { const n = mesh.nodeId[(i * mesh.ny + j) * mesh.nz + kt]; if (n >= 0) top.push(res.u[3 * n + 2]); }
  if (model.ndof <= 200000) {
    f64gz(res.u, name + '_u.f64.gz');
    f64gz(rhsCopy, name + '_rhs.f64.gz');
  } else {
    // large meshes: the bottom and top planes in full plus every 97th node, as (node, ux, uy, uz) rows
    const rows = [];
    const put = n => { rows.push(n, res.u[3 * n], res.u[3 * n + 1], res.u[3 * n + 2]); };
    for (let i = 0; i < mesh.nx; i++) for (let j = 0; j < mesh.ny; j++) for (const k of [0, spec.top_plane]) { const n = mesh.nodeId[(i * mesh.ny + j) * mesh.nz + k]; if (n >= 0) put(n); }
    for (let n = 0; n < mesh.nn; n += 97) put(n);
    f64gz(rows, name + '_usample.f64.gz');
    let s2 = 0; for (let q = 0; q < rhsCopy.length; q++) s2 += rhsCopy[q] * rhsCopy[q];
    var rhsNorm = Math.sqrt(s2);
  }
  const out = {
    name, nn: mesh.nn, ne: mesh.ne, ndof: model.ndof, constraints: list(mesh.constraints.dofs),
    iters: res.iters, relres: res.relres, converged: res.converged, tol: solver.opts.tol, ms,
    mg_levels: solver.pc && solver.pc.info ? solver.pc.info.levels : null,
    warpage: { signed: w.signed, mag: w.mag, sign: w.sign, plane: w.plane, rmin: w.rmin, rmax: w.rmax, nsites: w.ballW.length },
    top_w_max: top.reduce((a, b) => Math.max(a, b), -Infinity), top_w_min: top.reduce((a, b) => Math.min(a, b), Infinity),
    rhs_norm: typeof rhsNorm === 'number' ? rhsNorm : null,
  };
  fs.writeFileSync(path.join(HERE, name + '_golden.json'), JSON.stringify(out, null, 1));
  console.log(name, 'ndof', model.ndof, 'iters', res.iters, 'relres', res.relres.toExponential(3), 'ms', ms, 'warp', w.signed);
}

function elementCases() {
  const cases = [];
  const boxes = [[0.8333333333333357, 0.8333333333333357, 0.1], [0.1, 0.25, 0.05535714285714286], [1.0, 0.5, 2.0], [0.02, 0.5, 0.3]];
  const mats = [isoD(131000, 0.28, 0.45, new Float64Array(36)), isoD(5000, 0.3, 0.45, new Float64Array(36)), columnD(9000, 0.35, new Float64Array(36)), isoD(26000, 0.2, 0.45, new Float64Array(36))];
  const dF = [-125 * 2.6e-6, -125 * 40e-6, 100 * 22e-6, -90 * 15e-6];
  for (let q = 0; q < boxes.length; q++) {
    const [hx, hy, hz] = boxes[q], D = mats[q];
    const s0 = new Float64Array(6);
    for (let r = 0; r < 6; r++) s0[r] = (D[6 * r] + D[6 * r + 1] + D[6 * r + 2]) * dF[q];
    const M = shapeMatrices(null, D);
    const K = elementK33Fast(hx, hy, hz, M, new Float64Array(1089));
    const f = elementFth(hx, hy, hz, null, s0, new Float64Array(33));
    const c = condense33(K, f);
    cases.push({ kind: 'homog', hx, hy, hz, D: list(D), s0: list(s0), K33: list(K), f33: list(f), Kc: list(c.Kc), fc: list(c.fc), X: list(c.X), xf: list(c.xf) });
  }
  // layered element: three zeta-station weight matrices and thermal-stress weights
  {
    const hx = 0.7, hy = 0.4, hz = 0.3;
    const Ds = [isoD(26000, 0.2, 0.45, new Float64Array(36)), isoD(5000, 0.3, 0.45, new Float64Array(36)), isoD(17000, 0.25, 0.45, new Float64Array(36))];
    const wz = [0.2, 0.5, 0.3];
    const Wst = new Float64Array(108), Sst = new Float64Array(18);
    for (let j = 0; j < 3; j++) for (let q = 0; q < 36; q++) Wst[36 * j + q] = wz[j] * Ds[j][q];
    for (let j = 0; j < 3; j++) for (let r = 0; r < 6; r++) Sst[6 * j + r] = wz[j] * (Ds[j][6 * r] + Ds[j][6 * r + 1] + Ds[j][6 * r + 2]) * (-1e-3 * (j + 1));
    const M = shapeMatrices(Wst, null);
    const K = elementK33Fast(hx, hy, hz, M, new Float64Array(1089));
    const f = elementFth(hx, hy, hz, Sst, null, new Float64Array(33));
    const c = condense33(K, f);
    cases.push({ kind: 'layered', hx, hy, hz, W: list(Wst), V: list(Sst), K33: list(K), f33: list(f), Kc: list(c.Kc), fc: list(c.fc), X: list(c.X), xf: list(c.xf) });
  }
  fs.writeFileSync(path.join(HERE, 'element_golden.json'), JSON.stringify(cases));
  console.log('element cases', cases.length);
}

// Anand material point (src/core/13-fatigue.js:13-237) with the SAC305 constants of the
// viscoplastic cycling workload (s0 45.9 MPa, Q/R 7460 K, A 5.87e6 1/s, xi 2, m 0.0942,
// h0 9350 MPa, s_hat 58.3 MPa, n 0.015, a 1.5) and the SAC305 elastic table.
function anandCases() {
  const AnandPoint = get('AnandPoint'), anandSaturationStress = get('anandSaturationStress'), anandUniaxial = get('anandUniaxial');
  const mat = {
    id: 'sac305_cycling', name: 'SAC305 (cycling constants)',
    elastic: { type: 'iso', E: { type: 'table', T: [-40, 25, 75, 125], v: [57.4, 46.5, 38.2, 30.0], interp: 'linear', extrap: 'linear', u: 'GPa' }, nu: { v: 0.4 } },
    cte: { type: 'const', a: { v: 22.5 } },
    anand: { A: { v: 5.87e6 }, QR: { v: 7460 }, xi: { v: 2.0 }, m: { v: 0.0942 }, shat: { v: 58.3 }, n: { v: 0.015 }, h0: { v: 9350 }, a: { v: 1.5 }, s0: { v: 45.9 } },
  };
  const params = { A: 5.87e6, QR: 7460, xi: 2.0, m: 0.0942, shat: 58.3, n: 0.015, h0: 9350, a: 1.5, s0: 45.9, nu: 0.4 };
  const npt = 6, nstep = 36;
  const points = [];
  for (let p = 0; p < npt; p++) {
    const pt = new AnandPoint(mat);
    const out = { sig: new Float64Array(6), D: new Float64Array(36) };
    const steps = [];
    for (let k = 1; k <= nstep; k++) {
      const ph = 2 * Math.PI * k / 24;
      const T = 25 + 82.5 * Math.sin(ph + 0.3 * p) - 0.0;
      const dt = 20 + 10 * ((k + p) % 3);
      const amp = 2e-3 * (1 + 0.5 * p);
      const eps = Float64Array.from([amp * Math.sin(ph), -0.4 * amp * Math.sin(ph + 0.2), 0.3 * amp * Math.cos(ph * 0.7 + p),
        amp * 0.8 * Math.sin(ph + p), 0.5 * amp * Math.cos(ph), (p % 2 ? -1 : 1) * 0.6 * amp * Math.sin(2 * ph)]);
      const E = pt.E(T);
      const d = pt.step(eps, T, dt, out, true);
      pt.commit();
      steps.push({ eps: Array.from(eps), T, dt, E, d, sig: Array.from(out.sig), D: Array.from(out.D),
        ep: Array.from(pt.ep), s: pt.s, ebar: pt.ebar, W: pt.W, sigEq: pt.sigEq });
    }
    points.push(steps);
  }
  // V6 closed form (src/core/15-verify.js:237-248): saturation stress at constant strain rate
  const sat = [[25, 1e-3], [125, 1e-4], [-40, 1e-2]].map(([T, rate]) => ({ T, rate, closed: anandSaturationStress(mat, T, rate), uniaxial: anandUniaxial(mat, T, rate, 0.05, 400).sigFinal }));
  fs.writeFileSync(path.join(HERE, 'anand_golden.json'), JSON.stringify({ params, E_table: { T: [-40, 25, 75, 125], v: [57400, 46500, 38200, 30000] }, points, saturation: sat }));
  console.log('anand points', npt, 'steps', nstep, 'saturation', JSON.stringify(sat));
}

// Stress recovery on the bare-die mesh (src/core/09-assembly.js:480-572, src/core/11-analysis.js:264-368,
// src/core/12-post.js:180-309): element states and point stresses, the die top/bottom surface maps and the
// joint loads of a bump field, from the reference's own step solution.  The view and geometry objects are the
// minimal ones these functions read (view.model, view.T, view.point; geom.siParts; a bump field record).
function stressCases(name) {
  const spec = JSON.parse(fs.readFileSync(path.join(HERE, name + '_spec.json'), 'utf8'));
  const groups = spec.materials.map(m => ({ name: m.name, key: m.name, kind: 'homog', mat: m.name, stage: m.stage || 0, part: m.part, column: !!m.column }));
  const mesh = buildStructuredMesh(spec.xs, spec.ys, spec.zs, groups, assignFor(spec));
  mesh.constraints = constraints321(mesh, spec.constraint_span);
  if (spec.feature_x && spec.feature_x.length) { mesh.featureX = Int32Array.from(spec.feature_x); mesh.featureY = Int32Array.from(spec.feature_y); }
  mesh.groupIndex = {};  // as the package mesher records it (src/core/06-mesher.js:529-530)
  groups.forEach((g, i) => { mesh.groupIndex[g.key] = i; });
  const model = new FEModel(mesh, { materials: spec.materials.map(testMaterial) });
  const solver = new LinearSolver(model, { precond: 'mg', tol: 1e-12 });
  const rhs = new Float64Array(model.ndof);
  model.assembleStep(spec.Ta, spec.Tb, spec.stage || 0, rhs);
  const res = solver.solve(rhs, name);
  const rec = model.stepRecovery(spec.Ta, spec.Tb, spec.stage || 0);
  const cache = new Map();
  const state = e => { let st = cache.get(e); if (!st) { st = new Float64Array(model.stateSize(e)); model.accumulateStep(e, rec, res.u, st); cache.set(e, st); } return st; };
  const view = { model, T: spec.Tb, point: (e, xi, eta, zeta, out) => model.stateStress(e, state(e), xi, eta, zeta, out) };
  const dieSurfaceStress = get('dieSurfaceStress'), jointLoads = get('jointLoads'), PART = get('PART');
  const out = { name, T: spec.Tb, elements: [] };
  // element-level: states and stresses at three natural points of sampled elements
  const pts = [[-0.577, 0.577, 0.577], [0.3, -0.9, -0.2], [1, 1, -1]];
  const picks = spec.stress_elements || [0, Math.floor(mesh.ne / 3), Math.floor(mesh.ne / 2), mesh.ne - 1];
  for (const e of picks) {
    const st = state(e), o = { eps: new Float64Array(6), sig: new Float64Array(6) };
    const rows = pts.map(p => { model.stateStress(e, st, p[0], p[1], p[2], o); return { xez: p, sig: list(o.sig), eps: list(o.eps) }; });
    out.elements.push({ e, state: list(st), points: rows });
  }
  if (spec.stress_die) {
    const d = spec.stress_die;
    const geom = { siParts: [{ name: 'die', rect: { x0: d.rect[0], x1: d.rect[1], y0: d.rect[2], y1: d.rect[3] }, ztop: d.ztop, zb: d.zb, gkey: d.gkey, part: PART.DIE }] };
    for (const which of ['top', 'bottom']) {
      const r = dieSurfaceStress(view, geom, 0, which);
      out['die_' + which] = { nx: r.nx, ny: r.ny, xs: list(r.xs), ys: list(r.ys), smax: list(r.smax), sxx: list(r.sxx), syy: list(r.syy),
        peakInterior: r.peakInterior, peakX: r.peakX, peakY: r.peakY, gpPeakInterior: r.gpPeakInterior, gpPeakAll: r.gpPeakAll,
        corners: r.corners.map(c => c.value), midEdges: r.midEdges.map(c => c.value) };
    }
  }
  if (spec.stress_joints) {
    const j = spec.stress_joints;
    const d = { name: 'die', nbx: j.nbx, nby: j.nby, bx0: j.bx0, by0: j.by0, px: j.px, py: j.py, peripheral: false, rows: 0 };
    const r = jointLoads(view, d, j.zmid, j.part, j.A);
    out.joints = { N: list(r.N), V: list(r.V), szz: list(r.szz), worstN: r.worstN, worstV: r.worstV, worstNidx: r.worstNidx, worstVidx: r.worstVidx };
  }
  fs.writeFileSync(path.join(HERE, name + '_stress_golden.json'), JSON.stringify(out));
  console.log(name, 'stress golden: elements', out.elements.length, out.die_top ? 'die peak ' + out.die_top.peakInterior : '', out.joints ? 'worst N ' + out.joints.worstN : '');
}

// Warm start across the steps of a path (LinearSolver._warmStart/_storeSolution,
// src/core/10-solver.js:885-926): temperature-dependent moduli (every material but the
// die falls linearly from 1.2 E at -55 C to 0.6 E at 260 C) so that each step's matrix
// differs; a fresh preconditioner per step (the step index as the solve key), the default
// basis of 20 solutions.  Per step: the group data of prepareStep, the solve record and u.
function warmCases(name) {
  const spec = JSON.parse(fs.readFileSync(path.join(HERE, name + '_spec.json'), 'utf8'));
  const groups = spec.materials.map(m => ({ name: m.name, key: m.name, kind: 'homog', mat: m.name, stage: m.stage || 0, part: m.part, column: !!m.column }));
  const mesh = buildStructuredMesh(spec.xs, spec.ys, spec.zs, groups, assignFor(spec));
  mesh.constraints = constraints321(mesh, spec.constraint_span);
  if (spec.feature_x && spec.feature_x.length) { mesh.featureX = Int32Array.from(spec.feature_x); mesh.featureY = Int32Array.from(spec.feature_y); }
  const mats = spec.materials.map(m => {
    const t = testMaterial(m);
    if (m.part !== 4) t.elastic.E = { type: 'table', T: [-55, 260], v: [1.2 * m.E / 1000, 0.6 * m.E / 1000], interp: 'linear', extrap: 'linear', u: 'GPa' };
    return t;
  });
  const model = new FEModel(mesh, { materials: mats });
  const solver = new LinearSolver(model, { precond: 'mg', tol: 1e-12 });
  const temps = [150, 110, 70, 40, 25];
  const steps = [], us = [];
  for (let k = 0; k + 1 < temps.length; k++) {
    const rhs = new Float64Array(model.ndof);
    model.assembleStep(temps[k], temps[k + 1], 0, rhs);
    const gd = model.groups.map(g => ({ D: list(g._D), s0: list(g._s0), born: !!g._born }));
    const res = solver.solve(rhs, 'step ' + k, k);
    steps.push({ Ta: temps[k], Tb: temps[k + 1], groups: gd, iters: res.iters, iters0: res.iters0, relres: res.relres, warm: res.warm });
    us.push(...res.u);
  }
  fs.writeFileSync(path.join(HERE, name + '_warm_golden.json'), JSON.stringify({ name, tol: 1e-12, steps }));
  f64gz(us, name + '_warm_u.f64.gz');
  console.log(name, 'warm steps', steps.map(s => s.iters + '/' + s.iters0.toExponential(2)).join(' '));
}

// IC(0)-preconditioned solves (ICPC, src/core/10-solver.js:740-800, the Manteuffel
// shift schedule included) of the committed specs: iterations, shift, solution.
function icCases(name) {
  const spec = JSON.parse(fs.readFileSync(path.join(HERE, name + '_spec.json'), 'utf8'));
  const groups = spec.materials.map(m => ({ name: m.name, kind: 'homog', mat: m.name, stage: m.stage || 0, part: m.part, column: !!m.column }));
  const mesh = buildStructuredMesh(spec.xs, spec.ys, spec.zs, groups, assignFor(spec));
  mesh.constraints = constraints321(mesh, spec.constraint_span);
  const model = new FEModel(mesh, { materials: spec.materials.map(testMaterial) });
  const solver = new LinearSolver(model, { precond: 'ic', tol: 1e-10, warm: 0 });
  const rhs = new Float64Array(model.ndof);
  model.assembleStep(spec.Ta, spec.Tb, spec.stage || 0, rhs);
  const res = solver.solve(rhs, name);
  const out = { name, tol: 1e-10, iters: res.iters, relres: res.relres, converged: res.converged, precond: res.precond,
    shift: res.precond === 'ic' ? solver.pc.shift : null, fallback: res.fallback };
  fs.writeFileSync(path.join(HERE, name + '_ic_golden.json'), JSON.stringify(out));
  if (model.ndof <= 20000) f64gz(res.u, name + '_ic_u.f64.gz');
  console.log(name, 'IC(0): iters', res.iters, 'precond', res.precond, 'shift', out.shift, 'fallback', JSON.stringify(res.fallback), 'relres', res.relres.toExponential(2));
}

// The fallback to the other preconditioner after a failed run (LinearSolver.solve,
// src/core/10-solver.js:960-972), probed with small iteration caps.
function fallbackCases() {
  const out = [];
  for (const [name, maxit, precond] of [['plate_16', 50, 'ic'], ['plate_16', 5, 'mg'], ['cfg1', 40, 'ic']]) {
    const spec = JSON.parse(fs.readFileSync(path.join(HERE, name + '_spec.json'), 'utf8'));
    const groups = spec.materials.map(m => ({ name: m.name, kind: 'homog', mat: m.name, stage: m.stage || 0, part: m.part, column: !!m.column }));
    const mesh = buildStructuredMesh(spec.xs, spec.ys, spec.zs, groups, assignFor(spec));
    mesh.constraints = constraints321(mesh, spec.constraint_span);
    if (spec.feature_x && spec.feature_x.length) { mesh.featureX = Int32Array.from(spec.feature_x); mesh.featureY = Int32Array.from(spec.feature_y); }
    const model = new FEModel(mesh, { materials: spec.materials.map(testMaterial) });
    const solver = new LinearSolver(model, { precond, tol: 1e-10, warm: 0, maxit });
    const rhs = new Float64Array(model.ndof);
    model.assembleStep(spec.Ta, spec.Tb, spec.stage || 0, rhs);
    const res = solver.solve(rhs, name);
    out.push({ name, maxit, first: precond, iters: res.iters, relres: res.relres, converged: res.converged, precond: res.precond, fallback: res.fallback });
  }
  fs.writeFileSync(path.join(HERE, 'fallback_golden.json'), JSON.stringify(out, null, 1));
  console.log('fallback cases', JSON.stringify(out.map(o => [o.name, o.first, o.iters, o.precond])));
}

// Bump screening (screenJoints, cycleHistory, screeningSites: src/core/13-fatigue.js:
// 275-455): the reference's own loop over sites and cycle steps (AnandPoint.stepColumn /
// stepSeriesColumn) on a synthetic bump-layer strain table that stands in for
// siteStrainTable (the table is written out with the results), plus cycle histories and
// the site subsampling above the screening cap.
function screenCases() {
  const screenJoints = get('screenJoints'), cycleHistory = get('cycleHistory'), screeningSites = get('screeningSites');
  const solder = {
    id: 'sac305_cycling', name: 'SAC305 (cycling constants)',
    elastic: { type: 'iso', E: { type: 'table', T: [-40, 25, 75, 125], v: [57.4, 46.5, 38.2, 30.0], interp: 'linear', extrap: 'linear', u: 'GPa' }, nu: { v: 0.4 } },
    cte: { type: 'const', a: { v: 22.5 } },
    anand: { A: { v: 5.87e6 }, QR: { v: 7460 }, xi: { v: 2.0 }, m: { v: 0.0942 }, shat: { v: 58.3 }, n: { v: 0.015 }, h0: { v: 9350 }, a: { v: 1.5 }, s0: { v: 45.9 } },
  };
  const cu = { id: 'cu', name: 'Cu pillar', elastic: { type: 'iso', E: { type: 'table', T: [-55, 25, 150], v: [123, 120, 115], interp: 'linear', extrap: 'linear', u: 'GPa' }, nu: { v: 0.34 } }, cte: { type: 'const', a: { v: 17 } } };
  const synth = (sites, temps) => {
    const n = sites.x.length, nt = temps.length, out = new Float64Array(n * nt * 6);
    for (let q = 0; q < n; q++) {
      const x = sites.x[q], y = sites.y[q], miss = q % 37 === 5;
      for (let ti = 0; ti < nt; ti++) {
        const dT = temps[ti] - 25;
        const v = [1.6e-5 * dT * (1 + 0.2 * Math.sin(x)), 1.3e-5 * dT * (1 + 0.2 * Math.cos(y)), (2.4e-5 + 2e-6 * x) * dT,
          3e-6 * dT * Math.sin(x + y), 6e-6 * dT * (y - 1.5), 6e-6 * dT * (x - 2.0)];
        for (let c = 0; c < 6; c++) out[(q * nt + ti) * 6 + c] = miss ? NaN : v[c];
      }
    }
    return out;
  };
  let lastTable = null;
  globalThis.__screenTable = (sites, temps) => (lastTable = synth(sites, temps));
  get('siteStrainTable = function (an, dieIdx, sites, temps, layer) { return globalThis.__screenTable(sites, temps); }');
  const field = { name: 'die', nbx: 18, nby: 14, bx0: 0.3, by0: 0.25, px: 0.2, py: 0.2, peripheral: false, rows: 0 };
  const ring = { name: 'die', nbx: 20, nby: 16, bx0: 0.2, by0: 0.2, px: 0.18, py: 0.18, peripheral: true, rows: 2 };
  const defs = [
    { name: 'bump3', d: field, cycling: { Tmin: -40, Tmax: 125, ramp: 10, dwell: 15, cycles: 3 }, temps: [-40, 0, 25, 75, 125], pillar: null },
    { name: 'bump2', d: field, cycling: { Tmin: -55, Tmax: 125, ramp: 15, dwell: 10, cycles: 2 }, temps: [-55, -20, 25, 60, 125], pillar: null },
    { name: 'ddi2', d: field, cycling: { Tmin: -40, Tmax: 125, ramp: 10, dwell: 15, cycles: 2 }, temps: [-40, 25, 125], pillar: { mat: 'cu', fs: 0.35 } },
    { name: 'ring1', d: ring, cycling: { Tmin: 0, Tmax: 100, ramp: 20, dwell: 5, cycles: 1 }, temps: [0, 25, 100], pillar: null },
  ];
  const cases = [];
  for (const c of defs) {
    const an = { cfg: { materials: [solder, cu], cycling: c.cycling, bump: { solder: solder.id } }, temps: { cycling: c.temps }, mesh: { geom: {} } };
    const r = screenJoints(an, c.d, c.pillar ? { zmid: 0, part: 0 } : null, solder.id, c.pillar);
    f64gz(lastTable, 'screen_' + c.name + '_table.f64.gz');
    const nn = a => list(a).map(v => (isFinite(v) ? v : null));
    cases.push({ name: c.name, d: c.d, cycling: c.cycling, temps: c.temps, pillar: c.pillar, nSites: r.nSites, cycles: r.cycles,
      dW: nn(r.dW), dE: nn(r.dE), stab: nn(r.stab), top: r.top.map(t => t.q) });
    console.log('screen', c.name, 'sites', r.nSites, 'top dW', r.top[0].dW, 'ms', r.ms.toFixed(0));
  }
  const hists = [{ Tmin: -40, Tmax: 125, ramp: 10, dwell: 15, cycles: 3 }, { Tmin: -55, Tmax: 125, ramp: 15, dwell: 10, cycles: 2 },
    { Tmin: 0, Tmax: 100, ramp: 7, dwell: 5, cycles: 1 }].map(cy => { const h = cycleHistory(cy); return { cycling: cy, t: list(h.t), T: list(h.T), cycle: list(h.cycle), phase: h.phase }; });
  const big = { nbx: 160, nby: 150, bx0: 0.1, by0: 0.2, px: 0.15, py: 0.15, peripheral: false, rows: 0 };
  const per = { nbx: 60, nby: 50, bx0: 0.1, by0: 0.1, px: 0.15, py: 0.15, peripheral: true, rows: 5 };
  const subs = [[big, null], [per, 300]].map(([d, cap]) => { const r = screeningSites(d, cap); return { d, cap, subsampled: r.subsampled, n: r.sites.x.length, idx: r.idx ? list(r.idx) : null }; });
  fs.writeFileSync(path.join(HERE, 'screen_golden.json'), JSON.stringify({ solder: { params: { A: 5.87e6, QR: 7460, xi: 2.0, m: 0.0942, shat: 58.3, n: 0.015, h0: 9350, a: 1.5, s0: 45.9, nu: 0.4 }, E_table: { T: [-40, 25, 75, 125], v: [57400, 46500, 38200, 30000] }, cte: 22.5 },
    pillar_mat: { E_table: { T: [-55, 25, 150], v: [123000, 120000, 115000] }, nu: 0.34, cte: 17 }, cases, histories: hists, subsample: subs }));
}

// Bump submodel (BumpSubmodel, src/core/13-fatigue.js:514-1105): the reference's own global
// analysis of the bare-die preset at the Draft mesh (2 cycles) and one corner bump's
// submodel (8 x 8 x 8 elements). Written: the submodel's set-up data (planes, element
// materials, boundary nodes, the material grids of its elastic parts, the prescribed
// boundary displacement at every sample temperature of the cycle), the first steps' Newton
// records and displacement, and the result record.
function submodelCases() {
  const db = JSON.parse(fs.readFileSync(path.join(REF, 'src', 'materials-db.json'), 'utf8'));
  const presetConfig = get('presetConfig'), Analysis = get('Analysis'), BumpSubmodel = get('BumpSubmodel');
  const dieBumpSites = get('dieBumpSites'), materialTables = get('materialTables');
  const cfg = presetConfig(db, 'bare');
  cfg.mesh.preset = 'draft';
  cfg.cycling.cycles = 2;
  const an = new Analysis(cfg, { preset: 'draft', evaluations: 'full' });
  an.run();
  const d = an.mesh.geom.dies[0], g = an.mesh.geom;
  const sites = dieBumpSites(d);
  const q = 0, site = { x: sites.x[q], y: sites.y[q], ix: sites.ix[q], iy: sites.iy[q] };
  const sm = new BumpSubmodel(an, 0, site, { nxy: 8, nz: 8 });
  const m = sm.mesh;
  const cyc = cfg.cycling;
  const Tb = an.temps.all.filter(T => T >= cyc.Tmin - 1e-9 && T <= cyc.Tmax + 1e-9);
  const ub = new Float64Array(sm.ndof), bvals = [];
  for (const T of Tb) { sm.prescribe(T, ub); for (const n of sm.bNodes) bvals.push(ub[3 * n], ub[3 * n + 1], ub[3 * n + 2]); }
  const first = [];
  const orig = sm.stepTo.bind(sm);
  let u1 = null;
  sm.stepTo = (T, dt) => { const r = orig(T, dt); if (first.length < 6) first.push({ T, dt, converged: r.converged, newton: r.newton, dMax: r.dMax }); if (!u1 && r.converged) u1 = Float64Array.from(sm.u); return r; };
  const st = materialTables(sm.solderMat), ut = sm.ufTables;
  const a = sm.solderMat.anand, pv = x => (typeof x === 'object' ? x.v : x);
  const setup = {
    xs: list(m.xs), ys: list(m.ys), zs: list(m.zs), eGroup: list(m.eGroup), eIJK: list(m.eIJK), nn: m.nn, ne: m.ne,
    bNodes: list(sm.bNodes), zt: g.zt, zaf: g.zaf, die: { cx: d.cx, cy: d.cy, rt: d.profile.rt, rb: d.profile.rb }, site,
    solder: { A: pv(a.A), QR: pv(a.QR), xi: pv(a.xi), m: pv(a.m), shat: pv(a.shat), n: pv(a.n), h0: pv(a.h0), a: pv(a.a), s0: pv(a.s0), nu: st.nu },
    ufNu: ut.nu, fatigue: cfg.fatigue, cycling: cyc, tempsAll: an.temps.all, boundaryTemps: Tb,
  };
  f64gz(st.E, 'submodel_solderE.f64.gz'); f64gz(st.F, 'submodel_solderF.f64.gz');
  f64gz(ut.E, 'submodel_ufE.f64.gz'); f64gz(ut.F, 'submodel_ufF.f64.gz');
  f64gz(sm.subD, 'submodel_subD.f64.gz'); f64gz(sm.subF, 'submodel_subF.f64.gz');
  f64gz(sm.siD, 'submodel_siD.f64.gz'); f64gz(sm.siF, 'submodel_siF.f64.gz');
  f64gz(bvals, 'submodel_boundary.f64.gz');
  const t0 = Date.now();
  const r = sm.run();
  f64gz(u1, 'submodel_u1.f64.gz');
  const rec = { dW: r.dW, dWdie: r.dWdie, dWsub: r.dWsub, dE: r.dE, dEdie: r.dEdie, dEsub: r.dEsub, critical: r.critical,
    stabChange: r.stabChange, perCycle: r.perCycle, steps: r.steps, cutbacks: r.cutbacks, newtonPerStep: r.newtonPerStep,
    dJoint: r.dJoint, life: r.life, loop: r.loop, ndof: r.ndof, nElem: r.nElem,
    snaps: r.section.snaps.map(q => ({ T: q.T, t: q.t, cycle: q.cycle, phase: q.phase })) };
  const last = r.section.snaps[r.section.snaps.length - 1];
  fs.writeFileSync(path.join(HERE, 'submodel_section.f32.gz'), zlib.gzipSync(Buffer.from(Float32Array.from(last.sig).buffer), { level: 9 }));
  fs.writeFileSync(path.join(HERE, 'submodel_golden.json'), JSON.stringify({ setup, first, result: rec }));
  console.log('submodel', r.ndof, 'DOF; dW', r.dW, 'steps', r.steps, 'cutbacks', r.cutbacks, 'newton/step', r.newtonPerStep.toFixed(2), 'ms', Date.now() - t0);
}

// Thermal path planning and summation (temperatureSamples, thermalPathPoints,
// Analysis.evaluationTemperatures / buildPath / finishPath: src/core/11-analysis.js:29-261)
// on configuration records of several kinds (bare, lidded, stacked modules; full and preview
// evaluations; Tg windows): the reference's own methods called on plain objects that hold
// the fields they read.
function pathCases() {
  const temperatureSamples = get('temperatureSamples'), thermalPathPoints = get('thermalPathPoints'), Analysis = get('Analysis');
  const samples = [
    [-40, 125, [], [25]], [-55, 150, [120], [25]], [-40, 125, [105, 160], [25, 0]], [0, 100, [50.5], []],
  ].map(([a, b, tgs, ex]) => ({ args: [a, b, tgs, ex], out: temperatureSamples(a, b, tgs, ex) }));
  const points = [
    [217, 165, [], []], [165, 25, [120], [150, 217]], [25, -40, [], [-40, -30, 0, 10]], [25, 260, [160, 120], [100, 217, 218, 250]],
    [150, 150, [], []], [165, 25, [30], [25.5]],
  ].map(([a, b, tgs, st]) => ({ args: [a, b, tgs, st], out: thermalPathPoints(a, b, tgs, st) }));
  const plans = [];
  const cfgs = [
    { name: 'bare', packageType: 'bare', mods: false, tgs: [], stageTgs: { 0: [], 1: [], 2: [], 3: [] }, evaluations: 'full' },
    { name: 'lid', packageType: 'lid', mods: false, tgs: [160, 120], stageTgs: { 0: [160], 1: [160], 2: [160, 120], 3: [160, 120] }, evaluations: 'full' },
    { name: 'lid_sens', packageType: 'lid', mods: false, tgs: [160, 120], stageTgs: { 0: [160], 1: [160], 2: [160, 120], 3: [160, 120] }, evaluations: 'sensitivity' },
    { name: 'preview', packageType: 'stiffener', mods: false, tgs: [140], stageTgs: { 0: [140], 1: [140], 2: [140], 3: [140] }, evaluations: 'preview' },
    { name: 'modules_wcuf', packageType: 'lid', mods: 'wcuf', tgs: [150], stageTgs: { '-3': [], '-2': [150], '-1': [150], 0: [150], 1: [150], 2: [150], 3: [150] }, evaluations: 'full' },
    { name: 'modules_mold', packageType: 'bare', mods: 'mold', tgs: [], stageTgs: { '-3': [], '-2': [], '-1': [], 0: [], 1: [], 2: [], 3: [] }, evaluations: 'full' },
  ];
  for (const c of cfgs) {
    const cfg = {
      packageType: c.packageType, process: { Tsub: 217, Tjoin: 217, Tuf: 165, Tattach: 150 },
      reflow: { peak: 260, reportTemps: [25, 150, 220, 260] }, cycling: { Tmin: -40, Tmax: 125 },
      stack: { fill: c.mods === 'wcuf' ? 'wcuf' : 'mold', process: { Tddi: 240, Twcuf: 150, Tmold: 175 } },
    };
    const an = Object.create(Analysis.prototype);
    an.cfg = cfg; an.opts = { evaluations: c.evaluations };
    an.finalStage = c.packageType === 'bare' ? 2 : 3;
    an.hasModules = !!c.mods;
    an.tgs = c.tgs;
    an.stageTgs = new Map(Object.entries(c.stageTgs).map(([k, v]) => [Number(k), v]));
    an.solidi = c.mods ? [217, 221] : [217];
    const steps = an.buildPath();
    // summation with synthetic increments: du_k = (k + 1) * [1, -0.5, 0.25] + 1e-3 * [0, k, k * k]
    an.model = { ndof: 3 };
    an.du = steps.map((s, k) => Float64Array.from([(k + 1), -0.5 * (k + 1) + 1e-3 * k, 0.25 * (k + 1) + 1e-3 * k * k]));
    an.timings = {}; an.progress = () => {};
    an.finishPath();
    const fields = {};
    for (const [T, u] of an.fields) fields[T] = list(u);
    plans.push({ name: c.name, cfg, evaluations: c.evaluations, finalStage: an.finalStage, hasModules: an.hasModules, tgs: c.tgs, stageTgs: c.stageTgs,
      solidi: an.solidi, temps: an.temps, steps, path: { chain: an.path.chain, chainJoin: an.path.chainJoin, cpi: an.path.cpi, up: an.path.up, down: an.path.down,
        ddiJoin: an.path.ddiJoin, ddi: an.path.ddi }, fields, uCPI: an.uCPI ? list(an.uCPI) : null, uDDI: an.uDDI ? list(an.uDDI) : null });
    console.log('path', c.name, 'steps', steps.length, 'temps', an.temps.all.length);
  }
  fs.writeFileSync(path.join(HERE, 'path_golden.json'), JSON.stringify({ samples, points, plans }));
}

function resultsCases() {
  const encodeResultsFile = get('encodeResultsFile'), joinParts = get('joinParts'), decodeResultsFile = get('decodeResultsFile');
  const meshFingerprint = get('meshFingerprint'), resultsChunkSteps = get('resultsChunkSteps');
  const exportAnalysisState = get('exportAnalysisState'), Analysis = get('Analysis');
  const gz = (bytes, file) => fs.writeFileSync(path.join(HERE, file), zlib.gzipSync(Buffer.from(bytes), { level: 9 }));
  // 1. a value tree with every kind of value the format carries
  const shared = Float32Array.from([1.5, -2.25, 3e-8]);
  const tree = {
    cfg: { name: 'r\u00e9sum\u00e9 \u2713', packageType: 'lid', nested: [1, 0.1, -0, 1e21, 1.5e-7, 123456789012, 2 ** 53, -3.25e-300, 5e-324, 1e-6, 2.5e-7, 1e20],
      flags: [true, false, null], b: 1, 2: 'two', a: 'x', 1: 'one', quote: 'a"b\\c\n\t\u0001' },
    nums: [NaN, Infinity, -Infinity, 0],
    fields: new Map([[25, Float64Array.from([0.1, 0.2])], [-40, Float64Array.from([1, 2, 3])]]),
    temps: new Set([25, 150, 'x']),
    arrays: [Float64Array.from([Math.PI]), shared, shared, Int32Array.from([-1, 2]), Uint32Array.from([4294967295]),
      Int16Array.from([-32768]), Uint16Array.from([65535]), Int8Array.from([-128, 127]), Uint8Array.from([0, 255, 7])],
    empty: new Float32Array(0),
    skip: undefined, fn: () => 1,
  };
  let enc = encodeResultsFile(tree);
  gz(new Uint8Array(joinParts(enc.parts, enc.size)), 'results_tree.fcbgares.gz');
  // 2. mesh fingerprints of two golden meshes
  const fingerprints = {};
  for (const name of ['small_box', 'cfg1']) {
    const spec = JSON.parse(fs.readFileSync(path.join(HERE, name + '_spec.json'), 'utf8'));
    const groups = spec.materials.map(m => ({ name: m.name, kind: 'homog', mat: m.name, stage: m.stage || 0, part: m.part, column: !!m.column }));
    const mesh = buildStructuredMesh(spec.xs, spec.ys, spec.zs, groups, assignFor(spec));
    fingerprints[name] = { fingerprint: meshFingerprint(mesh), nn: mesh.nn, ne: mesh.ne };
  }
  // 3. an analysis state over the small_box mesh: the preview plan of a stiffener package and
  //    synthetic single-precision increments du_k[i] = fround((k + 1) * 0.001 * ((i % 97) - 48))
  const spec = JSON.parse(fs.readFileSync(path.join(HERE, 'small_box_spec.json'), 'utf8'));
  const groups = spec.materials.map(m => ({ name: m.name, kind: 'homog', mat: m.name, stage: m.stage || 0, part: m.part, column: !!m.column }));
  const mesh = buildStructuredMesh(spec.xs, spec.ys, spec.zs, groups, assignFor(spec));
  const cfg = { packageType: 'stiffener', process: { Tsub: 217, Tjoin: 217, Tuf: 165, Tattach: 150 },
    reflow: { peak: 260, reportTemps: [25, 150, 220, 260] }, cycling: { Tmin: -40, Tmax: 125 },
    stack: { fill: 'mold', process: { Tddi: 240, Twcuf: 150, Tmold: 175 } } };
  const an = Object.create(Analysis.prototype);
  an.cfg = cfg; an.opts = { preset: 'draft', evaluations: 'preview', precond: 'mg', hmin: 0.05 };
  an.finalStage = 3; an.hasModules = false; an.tgs = [140];
  an.stageTgs = new Map([[0, [140]], [1, [140]], [2, [140]], [3, [140]]]);
  an.solidi = [217];
  an.mesh = mesh;
  const ndof = 3 * mesh.nn;
  an.model = { ndof };
  const steps = an.buildPath();
  an.du = steps.map((s, k) => { const d = new Float32Array(ndof); for (let i = 0; i < ndof; i++) d[i] = Math.fround((k + 1) * 0.001 * ((i % 97) - 48)); return d; });
  const st = exportAnalysisState(an);
  enc = encodeResultsFile(st);
  const stBytes = new Uint8Array(joinParts(enc.parts, enc.size));
  gz(stBytes, 'results_state.fcbgares.gz');
  // restore: the decoded increments summed by finishPath
  const back = decodeResultsFile(stBytes.buffer);
  an.du = back.du; an.timings = {}; an.progress = () => {};
  an.finishPath();
  const fields = {};
  for (const [T, u] of an.fields) fields[T] = list(u);
  const out = { fingerprints, state: { cfg, opts: an.opts, ndof, nsteps: steps.length, fingerprint: st.fingerprint, plan: st.plan },
    restored: { fields, uCPI: an.uCPI ? list(an.uCPI) : null },
    chunkSteps: [1, 570, 76077, 3093159, 20000000].map(n => [n, resultsChunkSteps(n)]) };
  fs.writeFileSync(path.join(HERE, 'results_golden.json'), JSON.stringify(out));
  console.log('results: state', stBytes.length, 'bytes,', steps.length, 'steps; fingerprints', JSON.stringify(fingerprints));
}

// Reference-built package meshes (the drop-in's own inputs): the bare-die and the lidded
// default presets at the Draft mesh, built by the reference's buildPackageMesh (graded
// lines, fill fractions, layered substrate bands), with the FEModel's prepared data of two
// steps of its thermal path (the first chain step and the last, at 25 C in the final stage)
// and the reference LinearSolver's solution of each at tol 1e-12 with its JEITA warpage.
// Written per preset: package_<name>_golden.json.gz (mesh arrays, group and section step
// data, element sections, solutions and warpage).
function packageCases() {
  const db = JSON.parse(fs.readFileSync(path.join(REF, 'src', 'materials-db.json'), 'utf8'));
  const presetConfig = get('presetConfig'), Analysis = get('Analysis'), sectionWmean = get('sectionWmean');
  for (const [name, which] of [['bare', 'bare'], ['lidded', 'default']]) {
    const t0 = Date.now();
    const cfg = presetConfig(db, which);
    cfg.mesh.preset = 'draft';
    const an = new Analysis(cfg, { preset: 'draft', evaluations: 'full' });
    an.build();
    const steps = an.buildPath();
    const m = an.mesh, model = an.model;
    const eSection = new Int32Array(m.ne).fill(-1);
    for (let e = 0; e < m.ne; e++) if (model.groups[m.eGroup[e]].kind === 'layered') eSection[e] = model.sectionIndex(e);
    const chain = an.path.chain, pick = [chain[0], chain[chain.length - 1]];
    const solver = new LinearSolver(model, { precond: 'mg', tol: 1e-12 });
    const rhs = new Float64Array(model.ndof), W = new Float64Array(108);
    const out = [];
    for (const k of pick) {
      const s = steps[k];
      model.assembleStep(s.Ta, s.Tb, s.stage, rhs);
      const gD = [], gs0 = [], gborn = [];
      for (const g of model.groups) {
        if (g.kind === 'layered') { gD.push(...new Float64Array(36)); gs0.push(...new Float64Array(6)); gborn.push(1); continue; }
        gD.push(...g._D); gs0.push(...g._s0); gborn.push(g._born ? 1 : 0);
      }
      const sW = [], sV = [], sborn = [];
      for (const sec of model.sections) { sectionWmean(sec, s.Ta, s.Tb, W); sW.push(...W); sV.push(...sec._V); sborn.push(sec._born ? 1 : 0); }
      const rhsCopy = list(rhs);
      solver.basis = []; solver.pc = null;
      const res = solver.solve(rhs, 'package ' + name + ' step ' + k);
      const w = warpageJEITA(m, m.geom, res.u, cfg);
      out.push({ k, Ta: s.Ta, Tb: s.Tb, stage: s.stage, branch: s.branch, D: gD, s0: gs0, born: gborn, secW: sW, secV: sV,
        secBorn: sborn, u: list(res.u), rhs: rhsCopy, iters: res.iters, relres: res.relres, warpage: w.signed, warpMag: w.mag });
      console.log('package', name, 'step', k, s.Ta, '->', s.Tb, 'stage', s.stage, 'iters', res.iters, 'relres', res.relres, 'warpage', w.signed);
    }
    const bga = m.geom.bga;
    const rec = { name, preset: which, nx: m.xs.length, ny: m.ys.length, nz: m.zs.length, nn: m.nn, ne: m.ne,
      xs: list(m.xs), ys: list(m.ys), zs: list(m.zs), nodeId: list(m.nodeId), cellElems: list(m.cellElems),
      eGroup: list(m.eGroup), ePart: list(m.ePart), eScale: list(m.eScale), eSection: list(eSection),
      ngroups: model.groups.length, nsections: model.sections.length, groupKinds: model.groups.map(g => g.kind || 'homog'),
      constraints: list(m.constraints.dofs), featureX: list(m.featureX || []), featureY: list(m.featureY || []),
      bga: { x: list(bga.x), y: list(bga.y), zone: bga.zone || null }, steps: out, pathSteps: steps.length };
    fs.writeFileSync(path.join(HERE, 'package_' + name + '_golden.json.gz'), zlib.gzipSync(Buffer.from(JSON.stringify(rec)), { level: 9 }));
    console.log('package', name, m.nn * 3, 'DOF', m.ne, 'elements', model.sections.length, 'sections', Date.now() - t0, 'ms');
  }
}

// A stacked-module package (the reference's own V18 check, src/core/15-verify.js:577-585: the bare
// preset on a 14 mm substrate with one Si-base module of two chiplets, Draft mesh capped at 20000
// nodes). Its thermal path starts with the module stages (chiplet join, wafer underfill, mold cure;
// stage < 0), in which only the born module parts are assembled and every node they do not touch
// is held besides the module's 3-2-1 support (FEModel.assembleStep, src/core/09-assembly.js:392-450),
// and the multigrid aggregates follow the born parts (LinearSolver._buildPC, 10-solver.js:868-883).
// Written: package_modules_golden.json.gz with the FEModel's module constraints, the first step of
// every module stage, the first package step and the last chain step: prepared data, the held DOFs
// of the step, the reference's multigrid level sizes and its solution at tol 1e-12.
function modulesCase() {
  const db = JSON.parse(fs.readFileSync(path.join(REF, 'src', 'materials-db.json'), 'utf8'));
  const presetConfig = get('presetConfig'), Analysis = get('Analysis'), sectionWmean = get('sectionWmean');
  const defaultModule = get('defaultModule'), defaultChiplet = get('defaultChiplet');
  const t0 = Date.now();
  const cfg = presetConfig(db, 'bare');
  cfg.substrate.sx = cfg.substrate.sy = 14;
  const mod = defaultModule('M1', 0, 0, 'si');
  Object.assign(mod, { sx: 7, sy: 6 });
  mod.chiplets = [defaultChiplet('C1', -1.75, 0, 3.0, 5.0), defaultChiplet('C2', 1.75, 0, 3.0, 5.0)];
  cfg.dies = []; cfg.modules = [mod];
  const an = new Analysis(cfg, { preset: 'draft', evaluations: 'full', cap: 20000 });
  an.build();
  const steps = an.buildPath();
  const m = an.mesh, model = an.model;
  const eSection = new Int32Array(m.ne).fill(-1);
  for (let e = 0; e < m.ne; e++) if (model.groups[m.eGroup[e]].kind === 'layered') eSection[e] = model.sectionIndex(e);
  const chain = an.path.chain, pick = [];
  const seen = new Set();
  for (const k of chain) if (steps[k].stage < 0 && !seen.has(steps[k].stage)) { seen.add(steps[k].stage); pick.push(k); }
  pick.push(chain.find(k => steps[k].stage >= 0), chain[chain.length - 1]);
  const solver = new LinearSolver(model, { precond: 'mg', tol: 1e-12 });
  const rhs = new Float64Array(model.ndof), W = new Float64Array(108);
  const out = [];
  for (const k of pick) {
    const s = steps[k];
    model.assembleStep(s.Ta, s.Tb, s.stage, rhs);
    const gD = [], gs0 = [], gborn = [];
    for (const g of model.groups) {
      // a layered group's birth (FEModel.isBorn; prepareStep keeps no _born for it) scales its elements
      if (g.kind === 'layered') { gD.push(...new Float64Array(36)); gs0.push(...new Float64Array(6)); gborn.push(model.isBorn(g, s.stage) ? 1 : 0); continue; }
      gD.push(...g._D); gs0.push(...g._s0); gborn.push(g._born ? 1 : 0);
    }
    const sW = [], sV = [], sborn = [];
    for (const sec of model.sections) { sectionWmean(sec, s.Ta, s.Tb, W); sW.push(...W); sV.push(...sec._V); sborn.push(sec._born ? 1 : 0); }
    const held = s.stage < 0 ? list(model.moduleStageConstraints(s.stage)) : list(model.constraints);
    const rhsCopy = list(rhs);
    solver.basis = []; solver.pc = null;
    const res = solver.solve(rhs, 'modules step ' + k);
    const levels = solver.pc.info.levels.map(l => l.dof);
    out.push({ k, Ta: s.Ta, Tb: s.Tb, stage: s.stage, branch: s.branch, D: gD, s0: gs0, born: gborn, secW: sW, secV: sV,
      secBorn: sborn, held, levels, u: list(res.u), rhs: rhsCopy, iters: res.iters, relres: res.relres });
    console.log('modules step', k, s.Ta, '->', s.Tb, 'stage', s.stage, 'held', held.length, 'levels', levels.join(' '), 'iters', res.iters, 'relres', res.relres);
  }
  const bga = m.geom.bga;
  const rec = { name: 'modules', nx: m.xs.length, ny: m.ys.length, nz: m.zs.length, nn: m.nn, ne: m.ne,
    xs: list(m.xs), ys: list(m.ys), zs: list(m.zs), nodeId: list(m.nodeId), cellElems: list(m.cellElems),
    eGroup: list(m.eGroup), ePart: list(m.ePart), eScale: list(m.eScale), eSection: list(eSection),
    ngroups: model.groups.length, nsections: model.sections.length, groupKinds: model.groups.map(g => g.kind || 'homog'),
    groupStages: model.groups.map(g => g.stage || 0),
    constraints: list(m.constraints.dofs), moduleConstraints: list(model.moduleConstraints),
    featureX: list(m.featureX || []), featureY: list(m.featureY || []),
    bga: { x: list(bga.x), y: list(bga.y), zone: bga.zone || null }, steps: out, pathSteps: steps.length };
  fs.writeFileSync(path.join(HERE, 'package_modules_golden.json.gz'), zlib.gzipSync(Buffer.from(JSON.stringify(rec)), { level: 9 }));
  console.log('modules', m.nn * 3, 'DOF', m.ne, 'elements', model.sections.length, 'sections', Date.now() - t0, 'ms');
  // the whole thermal path of the same package as the reference's analysis runs it (preview
  // evaluations: module stages, then the package stages; warm starts along the path), with every
  // step's prepared data, the plan, the 25 C field and its JEITA warpage, and the chiplet-join state
  const t1 = Date.now();
  const ap = new Analysis(cfg, { preset: 'draft', evaluations: 'preview', cap: 20000 });
  ap.build();
  const psteps = ap.buildPath();
  const pm = ap.model, P = ap.path;
  ap.run();
  const stepData = psteps.map(s => {
    pm._prepKey = null;
    pm.prepareStep(s.Ta, s.Tb, s.stage);
    const gD = [], gs0 = [], gborn = [];
    for (const g of pm.groups) {
      if (g.kind === 'layered') { gD.push(...new Float64Array(36)); gs0.push(...new Float64Array(6)); gborn.push(pm.isBorn(g, s.stage) ? 1 : 0); continue; }
      gD.push(...g._D); gs0.push(...g._s0); gborn.push(g._born ? 1 : 0);
    }
    const sW = [], sV = [], sborn = [];
    for (const sec of pm.sections) { sectionWmean(sec, s.Ta, s.Tb, W); sW.push(...W); sV.push(...sec._V); sborn.push(sec._born ? 1 : 0); }
    return { Ta: s.Ta, Tb: s.Tb, stage: s.stage, branch: s.branch, D: gD, s0: gs0, born: gborn, secW: sW, secV: sV, secBorn: sborn };
  });
  const u25 = ap.fields.get(25), w25 = warpageJEITA(ap.mesh, ap.mesh.geom, u25, ap.cfg);
  const prec = { plan: { chain: P.chain, up: P.up, down: P.down, chainJoin: P.chainJoin, cpi: P.cpi, hasCPI: P.hasCPI,
    ddiJoin: P.ddiJoin, ddi: P.ddi, hasDDI: P.hasDDI }, temps: ap.temps.all, steps: stepData,
    iters: ap.solver.log.map(l => l.iters), u25: list(u25), warpage25: w25.signed, uDDI: ap.uDDI ? list(ap.uDDI) : null,
    same_mesh: ap.mesh.nn === m.nn && ap.mesh.ne === m.ne };
  fs.writeFileSync(path.join(HERE, 'package_modules_path.json.gz'), zlib.gzipSync(Buffer.from(JSON.stringify(prec)), { level: 9 }));
  console.log('modules path:', psteps.length, 'steps,', P.chain.filter(k => psteps[k].stage < 0).length, 'module steps in the chain, warpage',
    w25.signed, 'ddi', !!ap.uDDI, Date.now() - t1, 'ms');
}

// The reference's tornado (runSensitivityCase / tornadoFromCases, src/core/14-sensitivity.js:8-172)
// on the lidded default preset at the Draft mesh, for a set of its inputs spanning every kind
// (datasheet ranges, +-10 % material and geometry factors, +-10 C process and Tg temperatures):
// the base case and each input's low and high case run as the reference runs them (a full Draft
// analysis with the sensitivity evaluation set), with its outputs (warpRT, warpPeak, dieStress, dW)
// and the tornado table. For the device to run the same cases, each case keeps its thermal-path
// plan, its step data (per step: every group's D, s0, birth; every section's W, V, birth) and its
// mesh (cases whose mesh differs, the geometry inputs, carry their own). Step data blocks shared
// by several cases are stored once.
//   tornado_golden.json.gz  cases, plans, outputs, tornado rows, block indices, meshes
//   tornado_blocks.f64.gz   the distinct blocks (group: D 36, s0 6, born; section: W 108, V 18, born)
const TORNADO_IDS = ['coreCTE', 'coreE', 'abfA1', 'abfE', 'cuRatio', 'ufA1', 'ufE', 'ufTg', 'ufTsf', 'dieT', 'lidCTE',
  'lidT', 'adhE', 'Tattach', 'srA1', 'timE'];

function tornadoSetup() {
  const db = JSON.parse(fs.readFileSync(path.join(REF, 'src', 'materials-db.json'), 'utf8'));
  const cfg = get('presetConfig')(db, 'default');
  const inputs = get('sensitivityInputs')(cfg).filter(i => TORNADO_IDS.includes(i.id));
  const cases = [{ id: '__base', side: null, cfg }];
  for (const inp of inputs) {
    const b = get('sensitivityBounds')(cfg, inp);
    for (const side of ['lo', 'hi']) cases.push({ id: inp.id, side, value: b[side], base: b.base, cfg: get('applySensitivity')(cfg, inp, b[side]) });
  }
  return { cfg, inputs, cases };
}

// one case: the reference's runSensitivityCase, unrolled to keep the plan and step data
function tornadoCaseRecord(c) {
  const Analysis = get('Analysis'), sensitivityTemps = get('sensitivityTemps'), sensitivityOutputs = get('sensitivityOutputs');
  const sectionWmean = get('sectionWmean'), meshFingerprint = get('meshFingerprint');
  const an = new Analysis(c.cfg, { preset: 'draft', evaluations: 'sensitivity' });
  an.build();
  const temps = sensitivityTemps(c.cfg);
  an.evaluationTemperatures = () => temps;
  an.run();
  const outputs = sensitivityOutputs(an);
  const m = an.mesh, model = an.model, P = an.path, W = new Float64Array(108);
  const steps = [];
  for (const s of P.steps) {
    model.prepareStep(s.Ta, s.Tb, s.stage);
    const groups = model.groups.map(g => g.kind === 'layered' ? [...new Float64Array(42), model.isBorn(g, s.stage) ? 1 : 0]
      : [...g._D, ...g._s0, g._born ? 1 : 0]);
    const sections = model.sections.map(sec => { sectionWmean(sec, s.Ta, s.Tb, W); return [...W, ...sec._V, sec._born ? 1 : 0]; });
    steps.push({ Ta: s.Ta, Tb: s.Tb, stage: s.stage, branch: s.branch, groups, sections });
  }
  const eSection = new Int32Array(m.ne).fill(-1);
  for (let e = 0; e < m.ne; e++) if (model.groups[m.eGroup[e]].kind === 'layered') eSection[e] = model.sectionIndex(e);
  const g = m.geom;
  const mesh = { nx: m.xs.length, ny: m.ys.length, nz: m.zs.length, nn: m.nn, ne: m.ne,
    xs: list(m.xs), ys: list(m.ys), zs: list(m.zs), nodeId: list(m.nodeId), cellElems: list(m.cellElems),
    eGroup: list(m.eGroup), ePart: list(m.ePart), eScale: list(m.eScale), eSection: list(eSection),
    ngroups: model.groups.length, nsections: model.sections.length, groupKinds: model.groups.map(q => q.kind || 'homog'),
    groupKeys: model.groups.map(q => q.key), constraints: list(m.constraints.dofs),
    featureX: list(m.featureX || []), featureY: list(m.featureY || []),
    bga: { x: list(g.bga.x), y: list(g.bga.y), zone: g.bga.zone || null },
    siParts: g.siParts.map(p => ({ name: p.name, rect: p.rect, ztop: p.ztop, zb: p.zb, gkey: p.gkey, part: p.part })) };
  return { id: c.id, side: c.side, value: c.value, base: c.base, fingerprint: meshFingerprint(m), mesh, steps,
    chain: P.chain, up: P.up, down: P.down, temps, peak: c.cfg.reflow.peak, outputs,
    iters: an.solver.log.map(l => l.iters) };
}

function tornadoPart(i, n, file) {
  const { cases } = tornadoSetup();
  const out = [];
  for (let q = i; q < cases.length; q += n) {
    const t0 = Date.now();
    out.push(Object.assign({ index: q }, tornadoCaseRecord(cases[q])));
    console.log('tornado case', q, cases[q].id, cases[q].side || '', Date.now() - t0, 'ms');
  }
  fs.writeFileSync(file, JSON.stringify(out));
}

function tornadoCases(nproc) {
  const cp = require('child_process'), os = require('os');
  const { inputs, cases } = tornadoSetup();
  const tmp = fs.mkdtempSync(path.join(os.tmpdir(), 'tornado-'));
  const files = [];
  const kids = [];
  for (let i = 0; i < nproc; i++) {
    const f = path.join(tmp, 'part' + i + '.json');
    files.push(f);
    kids.push(new Promise((res, rej) => {
      const k = cp.spawn(process.execPath, [__filename, REF, 'tornado_part', String(i), String(nproc), f], { stdio: 'inherit' });
      k.on('exit', code => code === 0 ? res() : rej(new Error('tornado part ' + i + ' exited ' + code)));
    }));
  }
  return Promise.all(kids).then(() => {
    const recs = [];
    for (const f of files) recs.push(...JSON.parse(fs.readFileSync(f, 'utf8')));
    recs.sort((a, b) => a.index - b.index);
    // distinct meshes and step data blocks
    const meshes = {}, blocks = [], blockIdx = new Map();
    const put = arr => {
      const key = Buffer.from(Float64Array.from(arr).buffer).toString('base64');
      if (!blockIdx.has(key)) { blockIdx.set(key, blocks.length); blocks.push(arr); }
      return blockIdx.get(key);
    };
    const outCases = recs.map(r => {
      if (!meshes[r.fingerprint]) meshes[r.fingerprint] = r.mesh;
      const steps = r.steps.map(s => ({ Ta: s.Ta, Tb: s.Tb, stage: s.stage, branch: s.branch,
        groups: s.groups.map(put), sections: s.sections.map(put) }));
      return { id: r.id, side: r.side, value: r.value, base: r.base, mesh: r.fingerprint, steps, chain: r.chain, up: r.up,
        down: r.down, temps: r.temps, peak: r.peak, outputs: r.outputs, iters: r.iters };
    });
    const base = outCases[0].outputs, byInput = {};
    for (const inp of inputs) {
      const lo = outCases.find(c => c.id === inp.id && c.side === 'lo'), hi = outCases.find(c => c.id === inp.id && c.side === 'hi');
      byInput[inp.id] = { lo: lo.outputs, hi: hi.outputs, loValue: lo.value, hiValue: hi.value, baseValue: lo.base };
    }
    const tornado = get('tornadoFromCases')(inputs, base, byInput);
    const offs = [0];
    for (const b of blocks) offs.push(offs[offs.length - 1] + b.length);
    const flat = new Float64Array(offs[offs.length - 1]);
    blocks.forEach((b, q) => flat.set(b, offs[q]));
    f64gz(flat, 'tornado_blocks.f64.gz');
    const rec = { preset: 'default', inputs: inputs.map(i => ({ id: i.id, label: i.label, rule: i.rule, path: i.path })),
      cases: outCases, tornado, meshes, blockOffsets: offs };
    fs.writeFileSync(path.join(HERE, 'tornado_golden.json.gz'), zlib.gzipSync(Buffer.from(JSON.stringify(rec)), { level: 9 }));
    console.log('tornado:', outCases.length, 'cases,', Object.keys(meshes).length, 'meshes,', blocks.length, 'blocks,', flat.length, 'values');
    for (const k of Object.keys(tornado)) console.log(k, tornado[k].rows.map(r => r.id + ' ' + r.lo.toExponential(3) + ' ' + r.hi.toExponential(3)).join(' | '));
  });
}

// What the worst-bump dW output of each tornado case needs beside its solved path (screenBumps /
// screenJoints over the flip-chip bump layer, src/core/13-fatigue.js:359-364, 318-323): the bump
// layer's mid-plane and the dies' bump fields of the case's geometry, the cycling profile, and
// the solder's Anand constants and 1 C property tables (the same for every case here).
// Added to tornado_golden.json.gz by the argument tornado_screen (after tornado).
function tornadoScreen() {
  const Analysis = get('Analysis'), findMaterial = get('findMaterial'), AnandPoint = get('AnandPoint');
  const file = path.join(HERE, 'tornado_golden.json.gz');
  const rec = JSON.parse(zlib.gunzipSync(fs.readFileSync(file)).toString());
  const { cases } = tornadoSetup();
  let solder = null;
  rec.cases.forEach((rc, q) => {
    const c = cases[q];
    if (c.id !== rc.id || c.side !== rc.side) throw new Error('tornado case order changed');
    const an = new Analysis(c.cfg, { preset: 'draft', evaluations: 'sensitivity' });
    an.build();
    const g = an.mesh.geom;
    rc.screen = { zmid: g.zt + 0.5 * g.so, cycling: c.cfg.cycling,
      dies: g.dies.map(d => ({ nbx: d.nbx, nby: d.nby, bx0: d.bx0, by0: d.by0, px: d.px, py: d.py, peripheral: !!d.peripheral, rows: d.rows || 0 })) };
    const pt = new AnandPoint(findMaterial(c.cfg.materials, c.cfg.bump.solder));
    const s = { params: { A: pt.A, QR: pt.QR, xi: pt.xi, m: pt.m, shat: pt.shat, n: pt.n, h0: pt.h0, a: pt.a, s0: pt.s0, nu: pt.nu },
      E: list(pt.tables.E), F: list(pt.tables.F) };
    if (!solder) solder = s;
    else if (JSON.stringify(s) !== JSON.stringify(solder)) throw new Error('solder differs between tornado cases');
  });
  rec.solder = solder;
  fs.writeFileSync(file, zlib.gzipSync(Buffer.from(JSON.stringify(rec)), { level: 9 }));
  console.log('tornado screening data:', rec.cases.length, 'cases, dies per case', rec.cases[0].screen.dies.length);
}

// Step material data from the 1 C property grids (FEModel.prepareStep, src/core/09-assembly.js:346-365,
// with gridMean / gridStressIncrement, src/core/03-materials.js:370-409, and sectionWmean /
// sectionVstep, src/core/04-homogenize.js:206-221) on the lidded default preset: every group's
// grids (stiffness grid or constant, thermal strain grid, birth stage, kind) and every layered
// section's (weight grid, sub-layer stiffness and strain grids, Lagrange integrals, owner group),
// then the prepared data of every step of the full thermal path and of steps that probe the edges
// (a zero-length step, reversed and fractional steps, steps leaving the grid at either end).
//   prepare_grids.f64.gz   every grid, concatenated in the order prepare_golden.json lists
//   prepare_golden.json    the layout, the steps and their prepared data
function prepareCase() {
  const db = JSON.parse(fs.readFileSync(path.join(REF, 'src', 'materials-db.json'), 'utf8'));
  const Analysis = get('Analysis'), sectionWmean = get('sectionWmean');
  const cfg = get('presetConfig')(db, 'default');
  const an = new Analysis(cfg, { preset: 'draft', evaluations: 'full' });
  an.build();
  const steps = an.buildPath(), model = an.model;
  const flat = [];
  const put = arr => { const o = flat.length; for (let i = 0; i < arr.length; i++) flat.push(arr[i]); return [o, arr.length]; };
  const groups = model.groups.map(g => ({ kind: g.kind || 'homog', stage: g.stage,
    Dgrid: g.kind === 'layered' || g.Dconst ? null : put(g.Dgrid), Dconst: g.kind !== 'layered' && g.Dconst ? put(g.Dconst) : null,
    F: g.kind === 'layered' ? null : put(g.F) }));
  const sections = model.sections.map(s => ({ ns: s.ns, group: s.group == null ? -1 : s.group, W: put(s.W), D: put(s.D), F: put(s.F),
    Ls: put([].concat(...s.Ls.map(l => Array.from(l)))) }));
  const probes = steps.map(s => [s.Ta, s.Tb, s.stage]).concat([[25, 25, 3], [25, 260, 3], [260, 25, 2], [24.3, 187.7, 1],
    [187.7, 24.3, 0], [-95, -60, 3], [-60, -95, 3], [280, 320, 3], [-81.5, 301.25, 3], [150.5, 150.5, 1], [100, 99.999999, 2]]);
  const W = new Float64Array(108);
  const out = probes.map(([Ta, Tb, stage]) => {
    model._prepKey = null;
    model.prepareStep(Ta, Tb, stage);
    const D = [], s0 = [], born = [];
    for (const g of model.groups) {
      if (g.kind === 'layered') { D.push(...new Float64Array(36)); s0.push(...new Float64Array(6)); born.push(model.isBorn(g, stage) ? 1 : 0); continue; }
      D.push(...g._D); s0.push(...g._s0); born.push(g._born ? 1 : 0);
    }
    const secW = [], secV = [], secBorn = [];
    for (const sec of model.sections) { sectionWmean(sec, Ta, Tb, W); secW.push(...W); secV.push(...sec._V); secBorn.push(sec._born ? 1 : 0); }
    return { Ta, Tb, stage, D, s0, born, secW, secV, secBorn };
  });
  f64gz(flat, 'prepare_grids.f64.gz');
  fs.writeFileSync(path.join(HERE, 'prepare_golden.json'), JSON.stringify({ preset: 'default', tgrid: [-80, 300], groups, sections, steps: out }));
  console.log('prepare:', groups.length, 'groups,', sections.length, 'sections (sub-layers', sections.map(s => s.ns).join(' '), '),', flat.length, 'grid values,', out.length, 'steps');
}

const what = process.argv.slice(3);
if (what[0] === 'tornado_part') tornadoPart(+what[1], +what[2], what[3]);
if (what.includes('prepare')) prepareCase();
if (what.includes('tornado_screen')) tornadoScreen();
if (what.includes('tornado')) tornadoCases(+(what[what.indexOf('tornado') + 1]) || 4);
if (what.includes('package')) packageCases();
if (what.includes('package_modules')) modulesCase();
if (what.includes('ic')) { for (const n of ['small_box', 'plate_16', 'cfg1']) icCases(n); fallbackCases(); }
if (what.includes('warm')) for (const n of ['small_box', 'plate_16']) warmCases(n);
if (what.includes('stress')) for (const n of ['small_box', 'cfg1']) stressCases(n);
if (what.includes('stress_cfg2')) stressCases('cfg2');
if (what.includes('anand')) anandCases();
if (what.includes('screen')) screenCases();
if (what.includes('submodel')) submodelCases();
if (what.includes('path')) pathCases();
if (what.includes('results')) resultsCases();
if (!what.length || what.includes('element')) elementCases();
for (const n of ['small_box', 'plate_16', 'cfg1', 'cfg2']) if (!what.length || what.includes(n)) solveSpec(n);

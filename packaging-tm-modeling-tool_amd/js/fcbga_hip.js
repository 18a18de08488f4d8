'use strict';
// JS host side of the MI355X step solve: loads the N-API addon (fcbga_hip.node)
// and offers the reference's call surface for one thermal-path step.
//
//   const fh = require('./fcbga_hip.js');
//   const solver = new fh.HipStepSolver(mesh, {precision: 'mixed'});
//   const rec = solver.solveStep(fh.stepFromGroups(Ta, Tb, stage, groups), label);
//   // rec.u = du of the step, plus the LinearSolver.solve record fields
//   const view = new fh.HipStateView(solver, feModel, [{step, du: rec.u}], Tb);
//   dieSurfaceStress(view, geom, 0, 'top');   // the reference's post functions, unchanged
//
// install(Analysis) replaces Analysis.prototype.solveStep (src/core/11-analysis.js:189-206)
// so that assembleStep + LinearSolver.solve run on the GPU: the per-group step
// data come from FEModel.prepareStep (src/core/09-assembly.js:346-365), and a
// solve that does not converge throws the reference's message.  The desktop
// compute worker installs it in boot() when the addon loads; without it the
// JS path runs unchanged.
const path = require('path');

let addon = null;
function load() {
  if (!addon) addon = require(path.join(__dirname, '..', 'fcbga_hip.node'));
  return addon;
}

/**
 * Step data of a prepared FEModel step (FEModel.prepareStep, src/core/09-assembly.js:346-365): per
 * group the mean stiffness _D, the thermal stress increment _s0 and the birth flag; per layered
 * section (FEModel.sections, src/core/09-assembly.js:217-219) the zeta-station stiffness weights
 * averaged over the step (sectionWmean, src/core/04-homogenize.js:206, the core's own function,
 * passed in), its thermal stress weights _V and birth flag.  A layered group's elements read their
 * section's data; its own flag is the group's birth (FEModel.isBorn), which scales its elements.
 */
function stepFromModel(model, Ta, Tb, stage, sectionWmean) {
  const groups = model.groups, ng = groups.length, sections = model.sections || [], ns = sections.length;
  const D = new Float64Array(ng * 36), s0 = new Float64Array(ng * 6), born = new Uint8Array(ng);
  const isBorn = g => (typeof model.isBorn === 'function' ? model.isBorn(g, stage) : (g.stage || 0) <= stage);
  groups.forEach((g, q) => {
    if (g.kind === 'layered') { born[q] = isBorn(g) ? 1 : 0; return; }
    D.set(g._D, 36 * q);
    s0.set(g._s0, 6 * q);
    born[q] = g._born ? 1 : 0;
  });
  const step = { Ta, Tb, stage, D, s0, born };
  if (ns) {
    if (typeof sectionWmean !== 'function') throw new Error('layered sections need the core\'s sectionWmean');
    const secW = new Float64Array(ns * 108), secV = new Float64Array(ns * 18), secBorn = new Uint8Array(ns);
    sections.forEach((sec, q) => {
      sectionWmean(sec, Ta, Tb, secW.subarray(108 * q, 108 * q + 108));
      secV.set(sec._V, 18 * q);
      secBorn[q] = sec._born ? 1 : 0;
    });
    Object.assign(step, { secW, secV, secBorn });
  }
  return step;
}

/**
 * The mesh contract for createModel from a FEModel: its mesh plus group count, element sections and
 * the module constraints of stacked-module stages (FEModel.moduleConstraints, src/core/09-assembly.js:227).
 */
function meshForModel(mesh, model) {
  const sections = model.sections || [];
  const out = Object.assign({}, mesh, { ngroups: model.groups.length, nsections: sections.length });
  if (model.moduleConstraints) out.moduleConstraints = Int32Array.from(model.moduleConstraints);
  if (sections.length) {
    const eSection = new Int32Array(mesh.ne).fill(-1);
    for (let e = 0; e < mesh.ne; e++) if (model.groups[mesh.eGroup[e]].kind === 'layered') eSection[e] = model.sectionIndex(e);
    out.eSection = eSection;
  }
  return out;
}

/** Step data from per-group {D (36), s0 (6), born} records (the fields prepareStep sets: _D, _s0, _born). */
function stepFromGroups(Ta, Tb, stage, groups) {
  const ng = groups.length;
  const D = new Float64Array(ng * 36), s0 = new Float64Array(ng * 6), born = new Uint8Array(ng);
  groups.forEach((g, q) => {
    if (g.kind === 'layered') throw new Error('layered groups are not supported by the GPU step solve');
    D.set(g._D || g.D, 36 * q);
    s0.set(g._s0 || g.s0, 6 * q);
    born[q] = (g._born !== undefined ? g._born : g.born) ? 1 : 0;
  });
  return { Ta, Tb, stage, D, s0, born };
}

class HipStepSolver {
  constructor(mesh, opts) {
    this.hip = load();
    this.model = this.hip.createModel(mesh);
    this.opts = Object.assign({ tol: 1e-8, maxit: 2000, precision: 'mixed', precond: 'mg' }, opts || {});
    this.log = [];
  }
  /** extra: per-call options, e.g. {progress: fn(p), progressEvery: n} (live progress, in-solve cancel). */
  solveStep(step, label, extra) {
    const rec = this.hip.solveStep(this.model, step, extra ? Object.assign({}, this.opts, extra) : this.opts);
    rec.label = label;
    this.log.push({ label, iters: rec.iters, relres: rec.relres, converged: rec.converged, ms: rec.ms, pcMs: rec.pcMs, precond: rec.precond, fallback: rec.fallback });
    return rec;
  }
  /** One step of the stress histories of elements (accumulateStep, src/core/09-assembly.js:494-545). */
  stateAccumulate(step, du, elems, states, qOnly) {
    return this.hip.stateAccumulate(this.model, step, du, Int32Array.from(elems), states || null, !!qOnly);
  }
  /** Stress and strain at natural points of elements[rows] (stateStress, src/core/09-assembly.js:556-571). */
  stateStress(elems, states, rows, xez) {
    return this.hip.stateStress(this.model, Int32Array.from(elems), states, Int32Array.from(rows), Float64Array.from(xez));
  }
  /**
   * The thermal path kept on the device (Analysis.run / finishPath / restorePath,
   * src/core/11-analysis.js:212-261): pathSolve keeps step k's increment in device
   * memory, pathSum adds kept increments in the given order (a field, or the FP64
   * payload of a partial {k, du} message with ids = [k]), pathDu gives step k's
   * increment in single precision (what exportAnalysisState saves) and pathLoad
   * puts a saved one back in place of the solve.
   */
  pathSolve(k, step, label) {
    const rec = this.hip.pathCall(this.model, 'solve', k, step, this.opts);
    rec.label = label;
    this.log.push({ label, iters: rec.iters, relres: rec.relres, converged: rec.converged, ms: rec.ms, pcMs: rec.pcMs, precond: rec.precond, fallback: rec.fallback });
    return rec;
  }
  pathSum(ids) { return this.hip.pathCall(this.model, 'sum', Int32Array.from(ids)); }
  pathDu(k) { return this.hip.pathCall(this.model, 'du', k); }
  pathLoad(k, du) { this.hip.pathCall(this.model, 'load', k, du); }
  pathReset() { this.hip.pathCall(this.model, 'reset'); }
  destroy() {
    if (this.model) this.hip.destroyModel(this.model);
    this.model = null;
  }
}

const STATE = 81;

/**
 * Drop-in for the reference's StateView (src/core/11-analysis.js:336-365) in the
 * post-processing functions, which read view.model, view.T and view.point
 * (dieSurfaceStress, jointLoads, src/core/12-post.js:180-309): each element's
 * stress-history state is accumulated on the device over the steps that lead to
 * T ({step, du} in path order) and cached; prefetch(elems) batches the states of
 * many elements into one call per step.
 */
class HipStateView {
  constructor(solver, model, steps, T) {
    this.solver = solver; this.model = model; this.steps = steps; this.T = T;
    this.cache = new Map();
  }
  prefetch(elems) {
    const need = Array.from(new Set(Array.from(elems))).filter(e => !this.cache.has(e));
    if (!need.length) return;
    let st = null;
    for (const { step, du } of this.steps) st = this.solver.stateAccumulate(step, du, need, st);
    need.forEach((e, q) => this.cache.set(e, st.slice(q * STATE, (q + 1) * STATE)));
  }
  state(e) {
    if (!this.cache.has(e)) this.prefetch([e]);
    return this.cache.get(e);
  }
  point(e, xi, eta, zeta, out) {
    out = out || { eps: new Float64Array(6), sig: new Float64Array(6) };
    const r = this.solver.stateStress([e], this.state(e), [0], [xi, eta, zeta]);
    out.sig.set(r.sig); out.eps.set(r.eps);
    return out;
  }
}

/**
 * The sensitivity op with autograd.Function semantics over the addon pair:
 * forward(step, objective) solves the step and returns J = [warpage proxy on
 * plane kTop, von Mises p-norm over objective.vmElems], keeping the context on
 * the model; backward(gJ) returns the log-derivatives of gJ . J for every
 * group's E and CTE and every element's CTE from one adjoint solve (gJ of 3 nn
 * values instead: dJ/du of the caller's own objective of the step solution):
 * dJ/dE_g = dlnE[g] / E_g, dJ/dCTE_g = dlnCte[g] / CTE_g.  The reference's
 * finite-difference tornado (runSensitivityCase / tornadoFromCases,
 * src/core/14-sensitivity.js:8-172) runs beside it unchanged, each case's
 * solves going through install's solveStep.
 */
class HipSensitivity {
  constructor(solver) { this.solver = solver; this.saved = false; }
  forward(step, objective, opts) {
    const r = this.solver.hip.adjointForward(this.solver.model, step, objective, Object.assign({}, this.solver.opts, opts || {}));
    if (!r.converged) throw new Error('Linear solver did not converge (sensitivity forward): relative residual ' + r.relres.toExponential(2));
    this.saved = true;
    return r;
  }
  backward(gJ, elemCte) {
    if (!this.saved) throw new Error('sensitivity backward before forward');
    const r = this.solver.hip.adjointBackward(this.solver.model, Float64Array.from(gJ), elemCte !== false);
    if (!r.converged) throw new Error('Linear solver did not converge (sensitivity adjoint): relative residual ' + r.relres.toExponential(2));
    return r;
  }
}

/**
 * Bump screening with the joint loop on the device.  The reference's
 * screenJoints (src/core/13-fatigue.js:375-455) marches every joint through the
 * cycle on one thread; here the host only prepares what is shared by all joints
 * and hip.screenJoints runs one device thread per joint.  `core` supplies the
 * reference's own helpers: {screeningSites, siteStrainTable, findMaterial,
 * cycleHistory, AnandPoint, materialTables, gridAt, findInterval, CONST}.
 * The returned function has screenJoints' signature and record.
 */
function makeScreenJoints(core) {
  const hip = load();
  const PARAMS = ['A', 'QR', 'xi', 'm', 'shat', 'n', 'h0', 'a', 's0', 'nu'];

  // everything the device needs per history point, from the history and the
  // cycling temperatures of the strain table
  function historyArrays(hist, temps, point) {
    const nsteps = hist.T.length, f25 = point.thermalStrain(25);
    const arr = {
      nsteps, cycles: hist.cycles, T: Float64Array.from(hist.T), cycle: Int32Array.from(hist.cycle),
      dt: new Float64Array(nsteps), E: new Float64Array(nsteps), Fth: new Float64Array(nsteps),
      wt: new Float64Array(nsteps), wi: new Int32Array(nsteps), same: new Uint8Array(nsteps),
    };
    hist.T.forEach((T, k) => {
      const lo = core.findInterval(temps, T), span = temps[lo + 1] - temps[lo];
      arr.wi[k] = lo;
      arr.wt[k] = Math.min(1, Math.max(0, (T - temps[lo]) / span));
      arr.E[k] = point.E(T);
      arr.Fth[k] = point.thermalStrain(T) - f25;
      if (k === 0) return;
      arr.dt[k] = hist.t[k] - hist.t[k - 1];
      arr.same[k] = hist.phase[k] === hist.phase[k - 1] ? 1 : 0;
    });
    return arr;
  }

  // an elastic pillar below the solder cap: compliances (1/E, 1/G, 1/G) and
  // thermal strain from 25 C at every history point
  function pillarArrays(arr, tables) {
    const n = arr.nsteps, Cp = new Float64Array(3 * n), Fp = new Float64Array(n);
    const F25 = core.gridAt(tables.F, 25);
    for (let k = 0; k < n; k++) {
      const E = core.gridAt(tables.E, arr.T[k]), G = E / (2 + 2 * tables.nu);
      Cp.set([1 / E, 1 / G, 1 / G], 3 * k);
      Fp[k] = core.gridAt(tables.F, arr.T[k]) - F25;
    }
    arr.Cp = Cp;
    arr.Fp = Fp;
  }

  // indices of the finite entries of v, largest first
  function ranked(v) {
    const idx = [];
    for (let q = 0; q < v.length; q++) if (Number.isFinite(v[q])) idx.push(q);
    return idx.sort((p, q) => v[q] - v[p]);
  }

  return function screenJoints(an, d, layer, solderId, pillar) {
    const started = Date.now();
    const picked = core.screeningSites(d, pillar ? core.CONST.SCREEN_MAX_DDI : null);
    const sites = picked.sites, temps = an.temps.cycling;
    const table = Float64Array.from(core.siteStrainTable(an, null, sites, temps, layer));
    const point = new core.AnandPoint(core.findMaterial(an.cfg.materials, solderId));
    const hist = core.cycleHistory(an.cfg.cycling);
    const arr = historyArrays(hist, temps, point);
    arr.nsites = sites.x.length;
    arr.fs = 1;
    if (pillar) {
      arr.fs = pillar.fs;
      pillarArrays(arr, core.materialTables(core.findMaterial(an.cfg.materials, pillar.mat)));
    }
    const params = {};
    for (const k of PARAMS) params[k] = point[k];
    const res = hip.screenJoints(params, arr, table, temps.indexOf(25));
    const top = ranked(res.dW).slice(0, 20).map(q => ({ q, x: sites.x[q], y: sites.y[q], dW: res.dW[q], dE: res.dE[q] }));
    return {
      die: d.name, x: sites.x, y: sites.y, ix: sites.ix, iy: sites.iy, dW: res.dW, dE: res.dE, stab: res.stab,
      subsampled: picked.subsampled, nSites: sites.x.length, top, ms: Date.now() - started, cycles: hist.cycles,
    };
  };
}

/**
 * The device engine behind one of the reference's BumpSubmodel instances
 * (src/core/13-fatigue.js:514-1105).  The instance keeps its own run() (the
 * cycle, adaptive sub-stepping, cutbacks, hysteresis samples, layer averages
 * and the record); attachSubmodel replaces only the instance methods that do
 * the numerical work with device calls:
 *   stepTo(T, dt)          the Newton time step (:924-961): element pass with
 *                          the Anand points, tangent, IC(0)-PCG, commit
 *   _snapshot(), _restore  the committed state kept for a cutback (:1065-1075)
 *   _stressSnapshot(T)     corner stresses of every element (:1036-1062)
 * After each device step the committed displacement and material-point
 * states are written back into the instance (sm.u and the AnandPoint objects
 * of sm.anand), which is what run() reads.  The step material data come from
 * the instance's own _elastic, the boundary values from its prescribe.
 * Returns detach(), which frees the engine and puts the methods back.
 *   const detach = fh.attachSubmodel(sm, core);   // core: {CONST}
 *   try { rec = sm.run(); } finally { detach(); }
 */
function attachSubmodel(sm, core) {
  const hip = load();
  const msh = sm.mesh, nx = msh.xs.length, ny = msh.ys.length, nz = msh.zs.length;
  const ncell = (nx - 1) * (ny - 1) * (nz - 1), STATE_PT = 11;
  const cellOf = new Int32Array(msh.ne);
  for (let e = 0; e < msh.ne; e++) cellOf[e] = (msh.eIJK[3 * e] * (ny - 1) + msh.eIJK[3 * e + 1]) * (nz - 1) + msh.eIJK[3 * e + 2];
  // engine material codes (FCBGA_SUB_*) of the instance's groups
  const groupOfCode = [sm.MAT.SUB, sm.MAT.SOLDER, sm.MAT.UF, sm.MAT.SI];
  const cellMat = new Int8Array(ncell).fill(-1);
  for (let e = 0; e < msh.ne; e++) cellMat[cellOf[e]] = groupOfCode.indexOf(msh.eGroup[e]);
  if (cellMat.some(c => c < 0)) throw new Error('attachSubmodel: every cell needs one element of a known material');
  const probe = sm.anand.find(p => p);
  if (!probe) throw new Error('attachSubmodel: the submodel has no solder points');
  const params = {};
  for (const k of ['A', 'QR', 'xi', 'm', 'shat', 'n', 'h0', 'a', 's0', 'nu']) params[k] = probe[k];
  const f25 = probe.thermalStrain(25);
  const h = hip.submodelCreate({
    nx, ny, nz, xs: Float64Array.from(msh.xs), ys: Float64Array.from(msh.ys), zs: Float64Array.from(msh.zs),
    cellMat, cellLayer: new Uint8Array(ncell),
  }, params);
  const nb = sm.bNodes.length;
  if (hip.submodelCall(h, 'nboundary') !== nb) {
    hip.submodelCall(h, 'destroy');
    throw new Error('attachSubmodel: boundary node count differs from the engine\'s');
  }
  const scratch = new Float64Array(sm.ndof), ub = new Float64Array(3 * nb);
  const elasD = new Float64Array(4 * 36), elasSth = new Float64Array(4 * 6);

  // committed device state -> the instance
  const pull = () => {
    sm.u.set(hip.submodelCall(h, 'displacement'));
    if (sm.uN) sm.uN.set(sm.u);
    const st = hip.submodelCall(h, 'points');
    sm.anand.forEach((pt, idx) => {
      if (!pt) return;
      const o = (8 * cellOf[idx >> 3] + (idx & 7)) * STATE_PT;
      pt.ep.set(st.subarray(o, o + 6));
      pt.s = st[o + 6]; pt.ebar = st[o + 7]; pt.W = st[o + 8]; pt.sigEq = st[o + 9]; pt._lastD = st[o + 10];
    });
  };

  const saved = { stepTo: sm.stepTo, _snapshot: sm._snapshot, _restore: sm._restore, _stressSnapshot: sm._stressSnapshot };
  sm.stepTo = function (T, dt) {
    for (let c = 0; c < 4; c++)
      this._elastic(groupOfCode[c], this.Tn, T, elasD.subarray(36 * c, 36 * c + 36), elasSth.subarray(6 * c, 6 * c + 6));
    scratch.set(this.u);
    this.prescribe(T, scratch);
    this.bNodes.forEach((n, q) => ub.set(scratch.subarray(3 * n, 3 * n + 3), 3 * q));
    const r = hip.submodelCall(h, 'step', { T, dt, E: probe.E(T), fsolder: probe.thermalStrain(T) - f25, elasD, elasSth, ub });
    if (r.converged) {
      this.Tn = T;
      pull();
    }
    return { converged: r.converged, newton: r.newton, dMax: r.dMax };
  };
  sm._snapshot = function () {
    hip.submodelCall(h, 'save');
    return { Tn: this.Tn };
  };
  sm._restore = function (snap) {
    hip.submodelCall(h, 'restore');
    this.Tn = snap.Tn;
    pull();
  };
  sm._stressSnapshot = function (T) {
    const dev = hip.submodelCall(h, 'section', probe.E(T), probe.thermalStrain(T) - f25);
    const out = new Float32Array(48 * msh.ne);
    for (let e = 0; e < msh.ne; e++) out.set(dev.subarray(48 * cellOf[e], 48 * cellOf[e] + 48), 48 * e);
    return out;
  };
  return function detach() {
    Object.assign(sm, saved);
    hip.submodelCall(h, 'destroy');
  };
}

function install(Analysis, opts) {
  const original = Analysis.prototype.solveStep;
  const o = Object.assign({}, opts || {});
  const wmean = o.sectionWmean || (typeof globalThis.sectionWmean === 'function' ? globalThis.sectionWmean : null);
  delete o.sectionWmean;
  Analysis.prototype.solveStep = function (k) {
    const s = this.path.steps[k];
    // stacked-module stages (stage < 0) go to the device too: the library drops the unborn
    // elements, holds the module stage's constraint set and builds the stage's aggregates
    // from the module constraints passed with the mesh and the step's birth flags
    const sections = this.model.sections || [];
    if (sections.length && !wmean) return original.call(this, k);
    if (typeof this.checkCancel === 'function') this.checkCancel();
    if (!this._hip) {
      this._hip = new HipStepSolver(meshForModel(this.mesh, this.model),
        Object.assign({ tol: this.solver.opts.tol, maxit: this.solver.opts.maxit }, o));
    }
    this.model.prepareStep(s.Ta, s.Tb, s.stage);
    const label = 'step ' + k + ' (' + s.Ta + ' -> ' + s.Tb + ' C)';
    // live progress from inside the PCG (liveReport, src/core/10-solver.js:811-849, 1032) and the
    // in-solve cancel (checkCancel throws Error('cancelled'), which stops the device solve)
    const self = this, live = typeof globalThis.liveReport === 'function' ? globalThis.liveReport : null;
    const progress = p => {
      if (typeof self.checkCancel === 'function') self.checkCancel();
      if (!live) return;
      if (p.event === 'start') live({ event: 'start', label, n: 3 * self.mesh.nn, precond: p.precond, building: true, warm: 0 });
      else if (p.event === 'iter') live({ event: 'iter', label, iters: p.iters, it: p.iters, maxit: p.maxit, relres: p.relres, r0: p.r0, tol: p.tol, precond: p.precond }, true);
    };
    const res = this._hip.solveStep(stepFromModel(this.model, s.Ta, s.Tb, s.stage, wmean), label, { progress, progressEvery: 1 });
    res.T = s.Tb; res.stage = s.stage; res.step = k;
    if (this.timings) this.timings.solve = (this.timings.solve || 0) + res.ms;
    if (!res.converged) throw new Error('Linear solver did not converge (' + label + '): relative residual ' + res.relres.toExponential(2));
    return res;
  };
  return original;
}

module.exports = { load, HipStepSolver, HipStateView, HipSensitivity, makeScreenJoints, attachSubmodel, stepFromGroups, stepFromModel, meshForModel,
  install, STATE };

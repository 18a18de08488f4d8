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
  solveStep(step, label) {
    const rec = this.hip.solveStep(this.model, step, this.opts);
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
 * group's E and CTE and every element's CTE from one adjoint solve.  It takes
 * the place of the finite-difference tornado of the sensitivity study
 * (runSensitivityCase / tornadoFromCases, src/core/14-sensitivity.js:8-172):
 * dJ/dE_g = dlnE[g] / E_g, dJ/dCTE_g = dlnCte[g] / CTE_g.
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
 * screenJoints (src/core/13-fatigue.js:375-455) with the per-joint loop on the
 * device: the sites, the bump-layer strain table, the cycle history and the
 * per-step material data come from the reference's own functions, passed in
 * `core` ({screeningSites, siteStrainTable, findMaterial, cycleHistory,
 * AnandPoint, materialTables, gridAt, findInterval, clamp, CONST}); every
 * joint's AnandPoint.stepColumn / stepSeriesColumn march through the cycle runs
 * in one device thread (hip.screenJoints).  Returns the reference's record.
 */
function makeScreenJoints(core) {
  const hip = load();
  return function screenJoints(an, d, layer, solderId, pillar) {
    const t0 = Date.now();
    const { sites, subsampled } = core.screeningSites(d, pillar ? core.CONST.SCREEN_MAX_DDI : null);
    const temps = an.temps.cycling;
    const table = core.siteStrainTable(an, null, sites, temps, layer);
    const solder = core.findMaterial(an.cfg.materials, solderId);
    const hist = core.cycleHistory(an.cfg.cycling);
    const n = sites.x.length, nsteps = hist.t.length;
    const pt = new core.AnandPoint(solder);
    const st = {
      nsteps, cycles: hist.cycles, nsites: n, fs: pillar ? pillar.fs : 1,
      T: Float64Array.from(hist.T), dt: new Float64Array(nsteps), E: new Float64Array(nsteps), Fth: new Float64Array(nsteps),
      wt: new Float64Array(nsteps), wi: new Int32Array(nsteps), cycle: Int32Array.from(hist.cycle), same: new Uint8Array(nsteps),
    };
    for (let k = 0; k < nsteps; k++) {
      const i = core.findInterval(temps, hist.T[k]);
      st.wi[k] = i;
      st.wt[k] = core.clamp((hist.T[k] - temps[i]) / (temps[i + 1] - temps[i]), 0, 1);
      st.Fth[k] = pt.thermalStrain(hist.T[k]) - pt.thermalStrain(25);
      st.E[k] = pt.E(hist.T[k]);
      st.dt[k] = k > 0 ? hist.t[k] - hist.t[k - 1] : 0;
      st.same[k] = k > 0 && hist.phase[k] === hist.phase[k - 1] ? 1 : 0;
    }
    if (pillar) {
      const pm = core.findMaterial(an.cfg.materials, pillar.mat), tb = core.materialTables(pm);
      st.Cp = new Float64Array(3 * nsteps); st.Fp = new Float64Array(nsteps);
      for (let k = 0; k < nsteps; k++) {
        const Ep = core.gridAt(tb.E, hist.T[k]), Gp = Ep / (2 * (1 + tb.nu));
        st.Cp[3 * k] = 1 / Ep; st.Cp[3 * k + 1] = st.Cp[3 * k + 2] = 1 / Gp;
        st.Fp[k] = core.gridAt(tb.F, hist.T[k]) - core.gridAt(tb.F, 25);
      }
    }
    const params = { A: pt.A, QR: pt.QR, xi: pt.xi, m: pt.m, shat: pt.shat, n: pt.n, h0: pt.h0, a: pt.a, s0: pt.s0, nu: pt.nu };
    const r = hip.screenJoints(params, st, Float64Array.from(table), temps.indexOf(25));
    const order = Array.from({ length: n }, (_, q) => q).filter(q => isFinite(r.dW[q])).sort((a, b) => r.dW[b] - r.dW[a]);
    return { die: d.name, x: sites.x, y: sites.y, ix: sites.ix, iy: sites.iy, dW: r.dW, dE: r.dE, stab: r.stab, subsampled, nSites: n,
      top: order.slice(0, 20).map(q => ({ q, x: sites.x[q], y: sites.y[q], dW: r.dW[q], dE: r.dE[q] })), ms: Date.now() - t0, cycles: hist.cycles };
  };
}

/**
 * BumpSubmodel.run (src/core/13-fatigue.js:963-1090) with every time step on
 * the device engine (hip.submodelCreate / submodelCall): `sm` is the
 * reference's BumpSubmodel instance (its mesh, boundary, material tables and
 * global fields are set up by the reference's constructor), `core` the core's
 * {cycleHistory, AnandPoint, lifeEstimates, CONST}.  The step material data
 * come from the instance's own _elastic and the boundary from its prescribe;
 * the engine does stepTo, the commit, the cutback snapshot, the layer means,
 * the hysteresis samples and the section snapshots.  Returns the reference's
 * record.  Install with
 *   BumpSubmodel.prototype.run = function () { return fh.runSubmodel(this, core); };
 * (the homogenized-layer and elastic-solder check modes keep the JS run).
 */
function runSubmodel(sm, core) {
  const hip = load();
  const t0 = Date.now();
  const m = sm.mesh, MAT = sm.MAT, g = sm.an.mesh.geom, C = core.CONST;
  const nx = m.xs.length, ny = m.ys.length, nz = m.zs.length, ncell = (nx - 1) * (ny - 1) * (nz - 1);
  const cellMat = new Int8Array(ncell), cellLayer = new Uint8Array(ncell);
  const kind = [MAT.SUB, MAT.SOLDER, MAT.UF, MAT.SI];
  for (let e = 0; e < m.ne; e++) {
    const i = m.eIJK[3 * e], j = m.eIJK[3 * e + 1], k = m.eIJK[3 * e + 2], c = (i * (ny - 1) + j) * (nz - 1) + k;
    cellMat[c] = kind.indexOf(m.eGroup[e]);
    if (m.eGroup[e] === MAT.SOLDER) {
      const zc = 0.5 * (m.zs[k] + m.zs[k + 1]);
      cellLayer[c] = (zc > g.zaf - C.SUB_T_AVG ? 1 : 0) | (zc < g.zt + C.SUB_T_AVG ? 2 : 0);
    }
  }
  const pt = new core.AnandPoint(sm.solderMat);
  const h = hip.submodelCreate({ nx, ny, nz, xs: Float64Array.from(m.xs), ys: Float64Array.from(m.ys), zs: Float64Array.from(m.zs), cellMat, cellLayer },
    { A: pt.A, QR: pt.QR, xi: pt.xi, m: pt.m, shat: pt.shat, n: pt.n, h0: pt.h0, a: pt.a, s0: pt.s0, nu: pt.nu });
  try {
    const nb = sm.bNodes.length, u = new Float64Array(sm.ndof);
    let Tn = 25;
    const stepTo = (T, dt) => {
      const elasD = new Float64Array(144), elasSth = new Float64Array(24);
      for (let q = 0; q < 4; q++) sm._elastic(kind[q], Tn, T, elasD.subarray(36 * q, 36 * q + 36), elasSth.subarray(6 * q, 6 * q + 6));
      sm.prescribe(T, u);
      const ub = new Float64Array(3 * nb);
      for (let q = 0; q < nb; q++) { const n = sm.bNodes[q]; ub[3 * q] = u[3 * n]; ub[3 * q + 1] = u[3 * n + 1]; ub[3 * q + 2] = u[3 * n + 2]; }
      const r = hip.submodelCall(h, 'step', { T, dt, E: pt.E(T), fsolder: pt.thermalStrain(T) - pt.thermalStrain(25), elasD, elasSth, ub });
      if (r.converged) Tn = T;
      return r;
    };
    const hist = core.cycleHistory(sm.cfg.cycling), nsteps = hist.t.length, nc = hist.cycles;
    const phi = Math.atan2(sm.site.y - sm.die.cy, sm.site.x - sm.die.cx), cphi = Math.cos(phi), sphi = Math.sin(phi);
    const loop = { t: [], T: [], tau: [], gamma: [], tauSub: [], gammaSub: [] };
    const Wd = new Float64Array(nc + 1), Ws = new Float64Array(nc + 1), Ed = new Float64Array(nc + 1), Es = new Float64Array(nc + 1);
    const sample = (T, t) => {
      const s = hip.submodelCall(h, 'loopSample', T, pt.E(T), pt.thermalStrain(T) - pt.thermalStrain(25), cphi, sphi);
      loop.t.push(t); loop.T.push(T); loop.tau.push(s[0]); loop.gamma.push(s[1]); loop.tauSub.push(s[2]); loop.gammaSub.push(s[3]);
    };
    let totalNewton = 0, cutbacks = 0, steps = 0, nsubNext = 1;
    const snaps = [], snapOn = !sm.opts.uniformMaterial;
    sample(25, 0);
    for (let k = 1; k < nsteps; k++) {
      if (sm.opts.cancelled && sm.opts.cancelled()) throw new Error('cancelled');
      const T0 = hist.T[k - 1], T1 = hist.T[k], t0s = hist.t[k - 1], t1s = hist.t[k];
      let nsub = nsubNext, done = false, dLast = 0;
      for (let attempt = 0; attempt < 8 && !done; attempt++) {
        hip.submodelCall(h, 'save');
        const TnSave = Tn;
        let ok = true;
        for (let s = 1; s <= nsub; s++) {
          const res = stepTo(T0 + (T1 - T0) * s / nsub, (t1s - t0s) / nsub);
          totalNewton += res.newton; steps++;
          dLast = res.dMax;
          if (!res.converged || res.dMax > C.INELASTIC_STEP_MAX) { ok = false; break; }
        }
        if (ok) done = true;
        else { hip.submodelCall(h, 'restore'); Tn = TnSave; nsub *= 2; cutbacks++; }
      }
      nsubNext = dLast > 0.5 * C.INELASTIC_STEP_MAX ? Math.min(nsub * 2, 64) : (dLast < 0.2 * C.INELASTIC_STEP_MAX ? Math.max(1, nsub >> 1) : nsub);
      if (!done) throw new Error('Submodel step failed to converge after cutbacks at T = ' + T1.toFixed(1) + ' °C');
      sample(T1, t1s);
      const c = hist.cycle[k];
      if (snapOn && c === nc && (k === nsteps - 1 || hist.phase[k + 1] !== hist.phase[k]))
        snaps.push({ T: T1, t: t1s, cycle: c, phase: hist.phase[k], sig: hip.submodelCall(h, 'section', pt.E(T1), pt.thermalStrain(T1) - pt.thermalStrain(25)) });
      const lm = hip.submodelCall(h, 'layerMeans');
      Wd[c] = lm[0]; Ed[c] = lm[1]; Ws[c] = lm[2]; Es[c] = lm[3];
      if (sm.opts.progress) sm.opts.progress(k / (nsteps - 1), 'Submodel step ' + k + '/' + (nsteps - 1) + ' at ' + T1.toFixed(0) + ' °C');
    }
    const dWdie = Wd[nc] - Wd[nc - 1], dWsub = Ws[nc] - Ws[nc - 1], dEdie = Ed[nc] - Ed[nc - 1], dEsub = Es[nc] - Es[nc - 1];
    const dW = Math.max(dWdie, dWsub), dE = dWdie >= dWsub ? dEdie : dEsub;
    const prev = nc >= 2 ? Math.max(Wd[nc - 1] - Wd[nc - 2], Ws[nc - 1] - Ws[nc - 2]) : NaN;
    const stabChange = isFinite(prev) && prev > 0 ? Math.abs(dW - prev) / prev : NaN;
    const perCycle = [];
    for (let c = 1; c <= nc; c++) perCycle.push({ cycle: c, dWdie: Wd[c] - Wd[c - 1], dWsub: Ws[c] - Ws[c - 1] });
    const prof = sm.die.profile, dJoint = dWdie >= dWsub ? prof.rt * 2 : prof.rb * 2;
    return {
      die: sm.die.name, site: { x: sm.site.x, y: sm.site.y, ix: sm.site.ix, iy: sm.site.iy }, nxy: sm.opts.nxy, nz: sm.opts.nz, nElem: m.ne, ndof: sm.ndof,
      dW, dWdie, dWsub, dE, dEdie, dEsub, critical: dWdie >= dWsub ? 'die side' : 'substrate side', stabChange, stabWarn: isFinite(stabChange) && stabChange > C.STABILIZATION_WARN,
      perCycle, loop, life: core.lifeEstimates(dW, dE, sm.cfg.fatigue, dJoint), dJoint, steps, cutbacks, newtonPerStep: totalNewton / Math.max(steps, 1), ms: Date.now() - t0,
      section: snapOn ? {
        xs: Float64Array.from(m.xs), ys: Float64Array.from(m.ys), zs: Float64Array.from(m.zs), grp: Int8Array.from(m.eGroup), cx: sm.site.x, cy: sm.site.y, zt: g.zt, zaf: g.zaf,
        dieCx: sm.die.cx, dieCy: sm.die.cy, radii: { top: prof.rt, bottom: prof.rb, max: prof.rm },
        groups: [{ name: 'Substrate slice', material: 'top ' + Math.round(C.SUB_MARGIN * 1000) + ' µm, homogenized' }, { name: 'Solder', material: sm.solderMat.name },
          { name: 'Underfill', material: sm.ufMat ? sm.ufMat.name : '' }, { name: 'Silicon', material: '' }],
        snaps,
      } : null,
    };
  } finally {
    hip.submodelCall(h, 'destroy');
  }
}

function install(Analysis, opts) {
  const original = Analysis.prototype.solveStep;
  Analysis.prototype.solveStep = function (k) {
    if (this.model && this.model.sections && this.model.sections.length) return original.call(this, k);
    if (typeof this.checkCancel === 'function') this.checkCancel();
    if (!this._hip) {
      const mesh = Object.assign({}, this.mesh, { ngroups: this.model.groups.length });
      this._hip = new HipStepSolver(mesh, Object.assign({ tol: this.solver.opts.tol, maxit: this.solver.opts.maxit }, opts));
    }
    const s = this.path.steps[k];
    this.model.prepareStep(s.Ta, s.Tb, s.stage);
    const label = 'step ' + k + ' (' + s.Ta + ' -> ' + s.Tb + ' C)';
    const res = this._hip.solveStep(stepFromGroups(s.Ta, s.Tb, s.stage, this.model.groups), label);
    res.T = s.Tb; res.stage = s.stage; res.step = k;
    if (this.timings) this.timings.solve = (this.timings.solve || 0) + res.ms;
    if (!res.converged) throw new Error('Linear solver did not converge (' + label + '): relative residual ' + res.relres.toExponential(2));
    return res;
  };
  return original;
}

module.exports = { load, HipStepSolver, HipStateView, HipSensitivity, makeScreenJoints, runSubmodel, stepFromGroups, install, STATE };

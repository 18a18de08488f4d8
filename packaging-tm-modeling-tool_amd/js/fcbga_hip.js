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

module.exports = { load, HipStepSolver, HipStateView, HipSensitivity, makeScreenJoints, stepFromGroups, install, STATE };

"""CPU oracle of the global viscoplastic Newton (bump-array cycling workload).

TEST INFRASTRUCTURE ONLY: tests/ use it as the checker of the device engine
(fcbga_vp_*); the product never imports it.

It restates, with its own assembly and a direct sparse solve:
* the elastic elements through the C oracle's assembled system
  (oracle.System, the reference's element and assembly,
  src/core/09-assembly.js:26-165, 346-464) with the solder elements' scale
  set to zero;
* the solder elements: B-bar strain-displacement matrices (bbarMatrix,
  src/core/13-fatigue.js:474-490), Anand points at the 2 x 2 x 2 Gauss points
  (the C oracle's AnandPoint.step, src/core/13-fatigue.js:67-143, with the
  bracket-end guard of the global Newton: a non-positive effective stress is
  no flow), internal
  force sum w B^T sigma and tangent sum w B^T D B;
* the Newton time step of BumpSubmodel.stepTo (src/core/13-fatigue.js:924-961):
  residual tolerance relative to the force scale, increment tolerance, a
  final pass at the converged displacement, then the commit; here with an
  exact (direct) linear solve instead of the inexact PCG.
A point whose consistent tangent is not finite takes the secant tangent
K 1x1 + 2 G theta I_dev, theta = sigEq / sigma_trial >= 1e-3 (the rule of the
device engine; the reference never meets such points in its submodel).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import oracle as orc

GP = 1 / np.sqrt(3.0)
# corner order of a hex: (li, lj, lk) of corner a (src/core/06-mesher.js:46)
HEX = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])


def bbar(hx, hy, hz, xi, eta, zeta):
    """6 x 24 B-bar matrix of a box at (xi, eta, zeta): the normal rows take the
    volumetric part of the derivatives at the element centre."""
    s = 2 * HEX - 1

    def derivs(x, y, z):
        dx = s[:, 0] * (1 + s[:, 1] * y) * (1 + s[:, 2] * z) / 8 * (2 / hx)
        dy = s[:, 1] * (1 + s[:, 0] * x) * (1 + s[:, 2] * z) / 8 * (2 / hy)
        dz = s[:, 2] * (1 + s[:, 0] * x) * (1 + s[:, 1] * y) / 8 * (2 / hz)
        return dx, dy, dz

    dx, dy, dz = derivs(xi, eta, zeta)
    cx, cy, cz = derivs(0.0, 0.0, 0.0)
    B = np.zeros((6, 24))
    vol = np.stack([(cx - dx) / 3, (cy - dy) / 3, (cz - dz) / 3], axis=1)  # [8, 3]
    for a in range(8):
        c = 3 * a
        B[0:3, c:c + 3] = vol[a]
        B[0, c] += dx[a]
        B[1, c + 1] += dy[a]
        B[2, c + 2] += dz[a]
        B[3, c], B[3, c + 1] = dy[a], dx[a]
        B[4, c + 1], B[4, c + 2] = dz[a], dy[a]
        B[5, c], B[5, c + 2] = dz[a], dx[a]
    return B


def secant_tangent(E, nu, d, sigEq):
    G, K = E / (2 * (1 + nu)), E / (3 * (1 - 2 * nu))
    tri = sigEq + 3 * G * d
    th = sigEq / tri if tri > 0 else 1.0
    th = th if th >= 1e-3 else 1e-3
    D = np.zeros((6, 6))
    D[:3, :3] = K + 2 * G * th * (np.eye(3) - 1 / 3)
    D[3, 3] = D[4, 4] = D[5, 5] = G * th
    return D


class VpOracle:
    """Global viscoplastic Newton on a mesh whose group `solder["group"]` holds
    the Anand bumps; solder = dict(group, anand, nu, E_table, cte (ppm/K), Tref)."""

    def __init__(self, mesh, spec, solder):
        from fcbga_amd import meshspec  # the mesh/step-data helpers the tests share
        self.mesh, self.spec, self.sol = mesh, spec, solder
        self.ms = meshspec
        g = solder["group"]
        self.solder_elems = np.nonzero(mesh.e_group == g)[0]
        m2 = _copy_mesh(mesh)
        m2.e_scale = mesh.e_scale.copy()
        m2.e_scale[self.solder_elems] = 0.0
        m2.constraint_dofs = np.zeros(0, np.int32)  # unconstrained: prescribed values enter the free rows
        Tref = solder["Tref"]
        D, s0, born = meshspec.group_step_data(spec, Tref, Tref + 1.0, 0)
        S = orc.System(m2, D, s0, born)
        rp, col, val = S.csr()
        self.ndof = S.ndof
        self.K_el = _bcsr_to_csr(rp, col, val, S.nn)
        self.free = np.ones(self.ndof, bool)
        self.free[mesh.constraint_dofs] = False
        self.f_unit = S.rhs() * self.free  # thermal load of the elastic elements per kelvin from Tref
        self.bc = np.zeros(len(mesh.constraint_dofs))  # prescribed values of the constrained DOFs
        P = solder["anand"]
        self.P = [P[k] for k in ("A", "QR", "xi", "m", "shat", "n", "h0", "a", "s0")] + [solder["nu"]]
        self.nu = solder["nu"]
        ns = len(self.solder_elems)
        self.st = np.zeros((8 * ns, 11))
        self.st[:, 6] = P["s0"]
        # solder element geometry, DOFs and B-bar matrices at the 8 points
        self.dofs = np.empty((ns, 24), np.int64)
        self.B = np.empty((ns, 8, 6, 24))
        self.w = np.empty(ns)
        for q, e in enumerate(self.solder_elems):
            i, j, k = mesh.e_ijk[e]
            hx, hy, hz = mesh.xs[i + 1] - mesh.xs[i], mesh.ys[j + 1] - mesh.ys[j], mesh.zs[k + 1] - mesh.zs[k]
            self.w[q] = hx * hy * hz / 8
            self.dofs[q] = (3 * mesh.conn[e][:, None] + np.arange(3)[None, :]).reshape(-1)
            for gq in range(8):
                x, y, z = (2 * HEX[gq] - 1) * GP
                self.B[q, gq] = bbar(hx, hy, hz, x, y, z)
        self.u = np.zeros(self.ndof)

    def E(self, T):
        t = self.sol["E_table"]
        X, Y = np.asarray(t["T"], float), np.asarray(t["v"], float)
        i = int(np.clip(np.searchsorted(X, T, side="left") - 1, 0, len(X) - 2))
        return Y[i] + (Y[i + 1] - Y[i]) * (T - X[i]) / (X[i + 1] - X[i])

    def evaluate(self, u, T, dt, tangent=True, state=None):
        """Trial states from the committed ones at displacement u: residual
        r = f_th - K_el u - f_solder (zero on constrained DOFs), force scale,
        tangent K_t (CSR, constrained rows and columns removed but the
        diagonal) and the trial states."""
        st = self.st if state is None else state
        E = self.E(T)
        fsol = self.sol["cte"] * 1e-6 * (T - self.sol["Tref"])
        ue = u[self.dofs]  # [ns, 24]
        eps = np.einsum("qgrc,qc->qgr", self.B, ue).reshape(-1, 6)
        eps[:, :3] -= fsol
        tr, sig, D, d = orc.anand_batch(self.P, eps, T, dt, E, st, tangent=tangent, robust=True)
        sig = sig.reshape(-1, 8, 6)
        fe = self.w[:, None] * np.einsum("qgrc,qgr->qc", self.B, sig)
        fs = np.zeros(self.ndof)
        np.add.at(fs, self.dofs.reshape(-1), fe.reshape(-1))
        fth = self.f_unit * (T - self.sol["Tref"])
        ku = (self.K_el @ u) * self.free
        r = fth - ku - fs
        r[~self.free] = 0.0
        scale = np.linalg.norm(fth) + np.linalg.norm(fs) + np.linalg.norm(ku)
        Kt = None
        if tangent:
            Dq = D.reshape(-1, 36).copy()
            for p in np.nonzero(~np.isfinite(Dq).all(axis=1))[0]:
                Dq[p] = secant_tangent(E, self.nu, d[p], tr[p, 9]).reshape(-1)
            Dq = Dq.reshape(-1, 8, 6, 6)
            Ke = self.w[:, None, None] * np.einsum("qgra,qgrt,qgtb->qab", self.B, Dq, self.B)
            rows = np.repeat(self.dofs, 24, axis=1).reshape(-1)
            cols = np.tile(self.dofs, (1, 24)).reshape(-1)
            Ks = sp.csr_matrix((Ke.reshape(-1), (rows, cols)), shape=(self.ndof, self.ndof))
            Kt = self._constrain(self.K_el + Ks)
        return r, scale, Kt, tr, d

    def _constrain(self, K):
        """Rows and columns of constrained DOFs removed, unit diagonal there."""
        Pf = sp.diags(self.free.astype(float))
        return (Pf @ K @ Pf + sp.diags((~self.free).astype(float))).tocsr()

    def step(self, T, dt, maxit=25, res_tol=1e-6, du_tol=1e-8):
        """One time step from the committed state (the stepTo logic with exact
        inner solves); commits on convergence.  Returns the record."""
        u = self.u.copy()
        u[self.mesh.constraint_dofs] = self.bc
        conv, it = False, 0
        for it in range(1, maxit + 1):
            r, scale, Kt, _, _ = self.evaluate(u, T, dt)
            rn = np.linalg.norm(r)
            if rn <= res_tol * max(scale, 1e-300):
                conv = True
                break
            du = spla.spsolve(Kt.tocsc(), r)
            du[~self.free] = 0.0
            u += du
            if np.linalg.norm(du) < du_tol * max(np.linalg.norm(u), 1e-30):
                conv = True
                it += 1
                break
        dmax = 0.0
        if conv:
            _, _, _, tr, d = self.evaluate(u, T, dt, tangent=False)
            self.st[:, :10] = tr[:, :10]
            self.u = u
            dmax = float(d.max()) if len(d) else 0.0
        return {"converged": conv, "newton": min(it, maxit), "dmax": dmax}

    def linear_solution(self, T):
        """The elastic limit: K_lin u = f with the solder elements' B-bar
        stiffness at the elastic D and their thermal load (no flow)."""
        E, nu = self.E(T), self.nu
        G, K = E / (2 * (1 + nu)), E / (3 * (1 - 2 * nu))
        D = np.zeros((6, 6))
        D[:3, :3] = K + 2 * G * (np.eye(3) - 1 / 3)
        D[3, 3] = D[4, 4] = D[5, 5] = G
        fsol = self.sol["cte"] * 1e-6 * (T - self.sol["Tref"])
        sth = D @ np.array([fsol, fsol, fsol, 0, 0, 0])
        Ke = self.w[:, None, None] * np.einsum("qgra,rt,qgtb->qab", self.B, D, self.B)
        fe = self.w[:, None] * np.einsum("qgra,r->qa", self.B, sth)
        rows = np.repeat(self.dofs, 24, axis=1).reshape(-1)
        cols = np.tile(self.dofs, (1, 24)).reshape(-1)
        Ks = sp.csr_matrix((Ke.reshape(-1), (rows, cols)), shape=(self.ndof, self.ndof))
        f = self.f_unit * (T - self.sol["Tref"])
        np.add.at(f, self.dofs.reshape(-1), fe.reshape(-1))
        f[~self.free] = 0
        Kt = self._constrain(self.K_el + Ks)
        u = spla.spsolve(Kt.tocsc(), f)
        u[~self.free] = 0
        return u


def _copy_mesh(m):
    import copy
    return copy.copy(m)


def _bcsr_to_csr(rp, col, val, nn):
    rows, cols, vals = [], [], []
    brow = np.repeat(np.arange(nn), np.diff(rp))
    for p in range(3):
        for q in range(3):
            rows.append(3 * brow + p)
            cols.append(3 * col + q)
            vals.append(val[:, p, q])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(3 * nn, 3 * nn))

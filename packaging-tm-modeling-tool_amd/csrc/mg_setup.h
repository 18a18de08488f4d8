// mg_setup.h -- host-side symbolic set-up of the shell-aggregate multigrid
// hierarchy (aggregates, coarse-level patterns, in-plane Hermite transfer
// weights).  The numeric work (Galerkin products, factorisations, cycles) runs
// on the GPU; this file only decides the structure, which depends on the mesh
// and not on the step's stiffness values.
//
// Hierarchy of src/core/10-solver.js:106-119, 248-471:
//   level 0: the solid grid (3 DOFs per node);
//   level 1: shell aggregates, each a run of nodes of one (i, j) column split at
//            void gaps and at compliant layers, 6 DOFs (t_x, t_y, t_z, theta_x,
//            theta_y, eps_z) with u = t + theta x z' kinematics;
//   level 2+: in-plane semi-coarsening (every other grid line plus the last and
//            the feature lines), parents matched by z overlap, rigid-body height
//            shift, cubic Hermite deflection from (t_z, theta).
#pragma once

#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace mgs {

struct MeshView {
  int nx, ny, nz;
  const double *xs, *ys, *zs;
  const int32_t *node_id;     // nx*ny*nz (compact id or -1)
  const int32_t *cell_elems;  // 2 * ncells
  const int8_t *e_part;       // per element
  std::vector<int32_t> feature_x, feature_y;
  // stacked-module stage (buildShellLevel's act, src/core/10-solver.js:260-266): elements
  // and nodes (compact ids) in play; the others count as void.  Empty: all.
  std::vector<uint8_t> elem_on, node_on;
};

struct Level {
  int nx = 0, ny = 0;              // column grid of this level
  std::vector<double> xs, ys;      // column coordinates
  std::vector<int32_t> colStart;   // aggregates of column c: [colStart[c], colStart[c+1])
  std::vector<double> zref, zlo, zhi;
  std::vector<int32_t> keepX, keepY;
  int32_t nagg = 0;
  // matrix pattern (6x6 blocks), rows sorted
  std::vector<int32_t> rowPtr, col;
  // level 1 only: grid k-range of each aggregate (inclusive)
  std::vector<int32_t> k0, k1;
  // transfer from this level to the next coarser shell level (shell -> shell):
  // per aggregate f of this level, parents pIdx[pPtr[f]..pPtr[f+1]) with 6x6 weights pW
  std::vector<int32_t> pPtr, pIdx;
  std::vector<double> pW;
  // children (transpose) of the transfer INTO this level from the finer shell level
  std::vector<int32_t> chPtr, chIdx, chQ;
};

struct Hierarchy {
  std::vector<int32_t> nodeAgg;  // per grid node: level-1 aggregate or -1
  std::vector<double> zp;        // per grid node: z - zref of its aggregate
  std::vector<Level> L;          // shell levels 1..; L[0] is level 1
};

// parts that split a column into separate shell aggregates (src/core/10-solver.js:248)
inline bool compliant_part(int p) { return p == 1 || p == 2 || p == 3 || p == 5 || p == 7 || p == 9; }

// coarseLines (src/core/10-solver.js:171-177)
inline std::vector<int32_t> coarse_lines(int n, const std::vector<int32_t> &keep) {
  std::vector<char> in(n, 0);
  for (int i = 0; i < n; i += 2) in[i] = 1;
  in[n - 1] = 1;
  for (int32_t k : keep)
    if (k >= 0 && k < n) in[k] = 1;
  std::vector<int32_t> out;
  for (int i = 0; i < n; i++)
    if (in[i]) out.push_back(i);
  return out;
}

// Level-1 aggregates (buildShellLevel, src/core/10-solver.js:260-334) with
// mergeSingles 'below'.  A node column (i, j) classifies the layers around a
// node by the cell with linear index (i*(ny-1)+j)*(nz-1)+k, as the reference
// does; at j = ny-1 that index lands in the next row of cells and at i = nx-1
// past the end (read as no element).  Keeping this makes the hierarchy, and so
// the iteration counts, follow the reference's.
inline void build_level1(const MeshView &m, Hierarchy &H) {
  const int nx = m.nx, ny = m.ny, nz = m.nz, ncy = ny - 1, ncz = nz - 1;
  const int64_t ncells = (int64_t)(nx - 1) * ncy * ncz;
  const int32_t ncol = nx * ny;
  const int64_t G = (int64_t)ncol * nz;
  H.nodeAgg.assign(G, -1);
  H.zp.assign(G, 0.0);
  H.L.clear();
  H.L.emplace_back();
  Level &lv = H.L[0];
  lv.nx = nx; lv.ny = ny;
  lv.xs.assign(m.xs, m.xs + nx);
  lv.ys.assign(m.ys, m.ys + ny);
  lv.keepX = m.feature_x; lv.keepY = m.feature_y;
  lv.colStart.assign(ncol + 1, 0);
  const bool act = !m.elem_on.empty();
  auto cls = [&](int i, int j, int kk) -> int {
    if (kk < 0 || kk >= ncz) return 0;
    int64_t cc = ((int64_t)i * ncy + j) * ncz + kk;
    int any = 0, stiff = 0;
    for (int q = 0; q < 2; q++) {
      int32_t e = cc < ncells ? m.cell_elems[2 * cc + q] : -1;
      if (e < 0 || (act && !m.elem_on[e])) continue;
      any = 1;
      if (!compliant_part(m.e_part[e])) stiff = 1;
    }
    return !any ? 0 : (stiff ? 2 : 1);
  };
  std::vector<std::vector<int>> runs;
  int32_t na = 0;
  for (int32_t c = 0; c < ncol; c++) {
    lv.colStart[c] = na;
    int i = c / ny, j = c % ny;
    runs.clear();
    bool have = false;
    int prevK = -2, prevType = 0;
    for (int k = 0; k < nz; k++) {
      const int32_t nd = m.node_id[(int64_t)c * nz + k];
      if (nd < 0 || (act && !m.node_on[nd])) { have = false; prevK = -2; continue; }
      int below = cls(i, j, k - 1), above = cls(i, j, k);
      int type = (below == 2 || above == 2) ? 2 : 1;
      bool joined = false;
      if (have && prevK == k - 1 && below != 0) joined = below == 2 || (type == 1 && prevType == 1);
      if (!joined) { runs.emplace_back(); have = true; }
      runs.back().push_back(k);
      prevK = k; prevType = type;
    }
    for (int r = 0; r < (int)runs.size(); r++) {
      if (runs[r].size() != 1 || runs.size() < 2) continue;
      int t = r > 0 ? r - 1 : r + 1;
      int kk = runs[r][0];
      if (t < r) runs[t].push_back(kk); else runs[t].insert(runs[t].begin(), kk);
      runs.erase(runs.begin() + r);
      r--;
    }
    for (auto &run : runs) {
      double zs = 0;
      for (int k : run) { H.nodeAgg[(int64_t)c * nz + k] = na; zs += m.zs[k]; }
      lv.zlo.push_back(m.zs[run.front()]);
      lv.zhi.push_back(m.zs[run.back()]);
      double zr = zs / run.size();
      lv.zref.push_back(zr);
      lv.k0.push_back(run.front());
      lv.k1.push_back(run.back());
      for (int k : run) H.zp[(int64_t)c * nz + k] = m.zs[k] - zr;
      na++;
    }
  }
  lv.colStart[ncol] = na;
  lv.nagg = na;
  // pattern: aggregates of the nodes sharing an element with a node of the row
  lv.rowPtr.assign(na + 1, 0);
  std::vector<int32_t> mark(na, -1), tmp;
  std::vector<std::vector<int32_t>> rows(na);
  auto cell_has = [&](int ci, int cj, int ck) -> bool {
    if (ci < 0 || ci >= nx - 1 || cj < 0 || cj >= ncy || ck < 0 || ck >= ncz) return false;
    int64_t cc = ((int64_t)ci * ncy + cj) * ncz + ck;
    return m.cell_elems[2 * cc] >= 0 || m.cell_elems[2 * cc + 1] >= 0;
  };
  for (int32_t A = 0; A < na; A++) {
    tmp.clear();
    int32_t c = 0;
    // column of A: binary search in colStart
    c = (int32_t)(std::upper_bound(lv.colStart.begin(), lv.colStart.end(), A) - lv.colStart.begin()) - 1;
    int i = c / ny, j = c % ny;
    for (int k = lv.k0[A]; k <= lv.k1[A]; k++) {
      if (H.nodeAgg[(int64_t)c * nz + k] != A) continue;
      for (int di = -1; di <= 1; di++)
        for (int dj = -1; dj <= 1; dj++)
          for (int dk = -1; dk <= 1; dk++) {
            int ii = i + di, jj = j + dj, kk = k + dk;
            if (ii < 0 || ii >= nx || jj < 0 || jj >= ny || kk < 0 || kk >= nz) continue;
            int32_t B = H.nodeAgg[((int64_t)ii * ny + jj) * nz + kk];
            if (B < 0 || mark[B] == A) continue;
            // the two nodes share an element when one of the cells spanning both holds one
            bool share = false;
            for (int ci = std::max(i, ii) - 1; ci <= std::min(i, ii) && !share; ci++)
              for (int cj = std::max(j, jj) - 1; cj <= std::min(j, jj) && !share; cj++)
                for (int ck = std::max(k, kk) - 1; ck <= std::min(k, kk) && !share; ck++) share = cell_has(ci, cj, ck);
            if (!share) continue;
            mark[B] = A;
            tmp.push_back(B);
          }
    }
    std::sort(tmp.begin(), tmp.end());
    rows[A] = tmp;
    lv.rowPtr[A + 1] = lv.rowPtr[A] + (int32_t)tmp.size();
  }
  lv.col.reserve(lv.rowPtr[na]);
  for (auto &r : rows) lv.col.insert(lv.col.end(), r.begin(), r.end());
}

// hermitePair (src/core/10-solver.js:356-363)
inline void hermite_pair(double t, double h, bool alongY, double *WA, double *WB) {
  std::memset(WA, 0, 36 * sizeof(double));
  std::memset(WB, 0, 36 * sizeof(double));
  for (int d = 0; d < 6; d++) { WA[d * 6 + d] = 1 - t; WB[d * 6 + d] = t; }
  double h00 = 2 * t * t * t - 3 * t * t + 1, h10 = t * t * t - 2 * t * t + t, h01 = -2 * t * t * t + 3 * t * t,
         h11 = t * t * t - t * t;
  WA[14] = h00; WB[14] = h01;
  if (alongY) { WA[15] = h * h10; WB[15] = h * h11; }
  else { WA[16] = -h * h10; WB[16] = -h * h11; }
}

inline void mul66(const double *A, const double *B, double *C) {
  double t[36];
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) {
      double s = 0;
      for (int k = 0; k < 6; k++) s += A[i * 6 + k] * B[k * 6 + j];
      t[i * 6 + j] = s;
    }
  std::memcpy(C, t, sizeof t);
}

// rigid-body transfer from reference height zc to zf (src/core/10-solver.js:366-373)
inline void rbm_shift(double dz, double *M) {
  std::memset(M, 0, 36 * sizeof(double));
  for (int d = 0; d < 6; d++) M[d * 7] = 1;
  M[4] = dz; M[9] = -dz; M[17] = dz;
}

// Coarse shell level by in-plane semi-coarsening (buildCoarseShellLevel,
// src/core/10-solver.js:396-471).  Appends the new level and fills the finer
// level's transfer.
inline void build_coarse(Hierarchy &H) {
  Level &f = H.L.back();
  std::vector<int32_t> cx = coarse_lines(f.nx, f.keepX), cy = coarse_lines(f.ny, f.keepY);
  Level co;
  co.nx = (int)cx.size(); co.ny = (int)cy.size();
  for (int32_t I : cx) co.xs.push_back(f.xs[I]);
  for (int32_t J : cy) co.ys.push_back(f.ys[J]);
  co.colStart.assign(co.nx * co.ny + 1, 0);
  int32_t na = 0;
  for (int I = 0; I < co.nx; I++)
    for (int J = 0; J < co.ny; J++) {
      co.colStart[I * co.ny + J] = na;
      int32_t c = cx[I] * f.ny + cy[J];
      for (int32_t a = f.colStart[c]; a < f.colStart[c + 1]; a++) {
        co.zlo.push_back(f.zlo[a]); co.zhi.push_back(f.zhi[a]); co.zref.push_back(f.zref[a]);
        na++;
      }
    }
  co.colStart[co.nx * co.ny] = na;
  co.nagg = na;
  for (int32_t k : f.keepX) { auto it = std::find(cx.begin(), cx.end(), k); if (it != cx.end()) co.keepX.push_back((int32_t)(it - cx.begin())); }
  for (int32_t k : f.keepY) { auto it = std::find(cy.begin(), cy.end(), k); if (it != cy.end()) co.keepY.push_back((int32_t)(it - cy.begin())); }
  auto match = [&](int32_t C, double zlo, double zhi, double zref) -> int32_t {
    int32_t best = -1;
    double bo = 0;
    for (int32_t a = co.colStart[C]; a < co.colStart[C + 1]; a++) {
      double ov = std::min(zhi, co.zhi[a]) - std::max(zlo, co.zlo[a]);
      double score = ov > 0 ? ov + 1e-9 : ((zref >= co.zlo[a] - 1e-9 && zref <= co.zhi[a] + 1e-9) ? 1e-9 : -1);
      if (score > bo) { bo = score; best = a; }
    }
    return best;
  };
  struct Loc { int I0, I1; double t, h; };
  auto locate = [](const std::vector<int32_t> &lines, const std::vector<double> &coords, int i, double x) {
    Loc r;
    int I0 = 0;
    while (I0 + 1 < (int)lines.size() && lines[I0 + 1] <= i) I0++;
    r.I0 = I0;
    if (lines[I0] == i) { r.I1 = -1; r.t = 0; r.h = 0; return r; }
    r.I1 = I0 + 1;
    r.t = (x - coords[I0]) / (coords[I0 + 1] - coords[I0]);
    r.h = coords[I0 + 1] - coords[I0];
    return r;
  };
  f.pPtr.assign(f.nagg + 1, 0);
  f.pIdx.clear(); f.pW.clear();
  double WL[36], WR[36], WB[36], WT[36], tmp[36], sh[36], I6[36], out[36];
  std::memset(I6, 0, sizeof I6);
  for (int d = 0; d < 6; d++) I6[d * 7] = 1;
  for (int i = 0; i < f.nx; i++) {
    Loc px = locate(cx, co.xs, i, f.xs[i]);
    for (int j = 0; j < f.ny; j++) {
      Loc py = locate(cy, co.ys, j, f.ys[j]);
      int32_t c = i * f.ny + j;
      for (int32_t a = f.colStart[c]; a < f.colStart[c + 1]; a++) {
        int xC[2] = {px.I0, px.I1}, yC[2] = {py.I0, py.I1};
        int nxC = px.I1 < 0 ? 1 : 2, nyC = py.I1 < 0 ? 1 : 2;
        int32_t par[2][2] = {{-1, -1}, {-1, -1}};
        bool complete = true;
        for (int u = 0; u < nxC; u++)
          for (int v = 0; v < nyC; v++) {
            par[u][v] = match(xC[u] * co.ny + yC[v], f.zlo[a], f.zhi[a], f.zref[a]);
            if (par[u][v] < 0) complete = false;
          }
        auto push = [&](int u, int v, const double *W) {
          int32_t ag = par[u][v];
          rbm_shift(f.zref[a] - co.zref[ag], sh);
          mul66(W, sh, out);
          f.pIdx.push_back(ag);
          f.pW.insert(f.pW.end(), out, out + 36);
        };
        size_t before = f.pIdx.size();
        if (complete) {
          if (px.I1 >= 0) hermite_pair(px.t, px.h, false, WL, WR);
          if (py.I1 >= 0) hermite_pair(py.t, py.h, true, WB, WT);
          if (px.I1 < 0 && py.I1 < 0) push(0, 0, I6);
          else if (py.I1 < 0) { push(0, 0, WL); push(1, 0, WR); }
          else if (px.I1 < 0) { push(0, 0, WB); push(0, 1, WT); }
          else {
            mul66(WB, WL, tmp); push(0, 0, tmp);
            mul66(WB, WR, tmp); push(1, 0, tmp);
            mul66(WT, WL, tmp); push(0, 1, tmp);
            mul66(WT, WR, tmp); push(1, 1, tmp);
          }
        } else {
          double wx[2] = {px.I1 < 0 ? 1 : 1 - px.t, px.t}, wy[2] = {py.I1 < 0 ? 1 : 1 - py.t, py.t};
          double sum = 0;
          for (int u = 0; u < nxC; u++)
            for (int v = 0; v < nyC; v++)
              if (par[u][v] >= 0 && wx[u] * wy[v] > 0) sum += wx[u] * wy[v];
          for (int u = 0; u < nxC; u++)
            for (int v = 0; v < nyC; v++) {
              double w = wx[u] * wy[v];
              if (par[u][v] >= 0 && w > 0) {
                double Wd[36];
                std::memset(Wd, 0, sizeof Wd);
                for (int d = 0; d < 6; d++) Wd[d * 7] = w / sum;
                push(u, v, Wd);
              }
            }
        }
        f.pPtr[a + 1] = (int32_t)(f.pIdx.size() - before);
      }
    }
  }
  for (int32_t a = 0; a < f.nagg; a++) f.pPtr[a + 1] += f.pPtr[a];
  // children lists of the coarse aggregates
  co.chPtr.assign(na + 1, 0);
  for (int32_t q = 0; q < (int32_t)f.pIdx.size(); q++) co.chPtr[f.pIdx[q] + 1]++;
  for (int32_t I = 0; I < na; I++) co.chPtr[I + 1] += co.chPtr[I];
  co.chIdx.assign(f.pIdx.size(), 0);
  co.chQ.assign(f.pIdx.size(), 0);
  std::vector<int32_t> fill(co.chPtr.begin(), co.chPtr.end() - 1);
  for (int32_t a = 0; a < f.nagg; a++)
    for (int32_t q = f.pPtr[a]; q < f.pPtr[a + 1]; q++) {
      int32_t I = f.pIdx[q];
      co.chIdx[fill[I]] = a;
      co.chQ[fill[I]] = q;
      fill[I]++;
    }
  // coarse pattern: parents of the two ends of every fine block
  co.rowPtr.assign(na + 1, 0);
  std::vector<int32_t> mark(na, -1), tmpr;
  std::vector<int32_t> cols;
  for (int32_t I = 0; I < na; I++) {
    tmpr.clear();
    for (int32_t q = co.chPtr[I]; q < co.chPtr[I + 1]; q++) {
      int32_t a = co.chIdx[q];
      for (int32_t k = f.rowPtr[a]; k < f.rowPtr[a + 1]; k++) {
        int32_t b = f.col[k];
        for (int32_t p = f.pPtr[b]; p < f.pPtr[b + 1]; p++) {
          int32_t J = f.pIdx[p];
          if (mark[J] != I) { mark[J] = I; tmpr.push_back(J); }
        }
      }
    }
    std::sort(tmpr.begin(), tmpr.end());
    co.rowPtr[I + 1] = co.rowPtr[I] + (int32_t)tmpr.size();
    cols.insert(cols.end(), tmpr.begin(), tmpr.end());
  }
  co.col = cols;
  H.L.push_back(std::move(co));
}

// Whole hierarchy with the reference's stopping rules (src/core/10-solver.js:632-645,
// src/core/00-constants.js:79-81).
inline void build_hierarchy(const MeshView &m, Hierarchy &H) {
  build_level1(m, H);
  for (int guard = 0; guard < 20; guard++) {
    Level &lv = H.L.back();
    int32_t ncol = lv.nx * lv.ny, ndof = 6 * lv.nagg;
    if (ncol <= 9 || ndof <= 900 || (lv.nx <= 2 && lv.ny <= 2)) break;
    size_t ncx = coarse_lines(lv.nx, lv.keepX).size(), ncy = coarse_lines(lv.ny, lv.keepY).size();
    if ((double)ncx * ncy > 0.7 * ncol) break;
    build_coarse(H);
  }
}

}  // namespace mgs

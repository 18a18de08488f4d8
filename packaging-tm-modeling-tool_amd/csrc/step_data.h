// step_data.h -- the step material data of the thermal path from 1 degC
// property grids (host only): per group the mean stiffness over the step and
// the thermal stress increment, per layered section the mean zeta-station
// stiffness weights and the thermal stress weights.  What FEModel.prepareStep
// hands to assembly (src/core/09-assembly.js:346-365).
//
// A grid holds `stride` values per degree C from TGRID_MIN to TGRID_MAX and is
// linear between the points (clamped outside).  A step [Ta, Tb] (either order)
// is cut at the grid points; on each piece the stiffness is linear in T and
// the thermal strain's derivative constant, so the midpoint value times the
// piece's strain change integrates D dF exactly, and the length-weighted
// midpoint values give the exact mean (src/core/03-materials.js:364-409).
#pragma once

#include <stdint.h>

#include <algorithm>
#include <cmath>

namespace sd {

constexpr int TGRID_MIN = -80;
constexpr int TGRID_N = 381;  // -80 .. 300 degC

// value of component c at T (gridAt, src/core/03-materials.js:347-355): the end
// values outside the grid
inline double grid_value(const double *g, double T, int stride, int c) {
  const double x = T - TGRID_MIN;
  if (x <= 0) return g[c];
  if (x >= TGRID_N - 1) return g[(int64_t)(TGRID_N - 1) * stride + c];
  const int i = (int)std::floor(x);
  const double t = x - i;
  const double a = g[(int64_t)i * stride + c], b = g[(int64_t)(i + 1) * stride + c];
  return a + t * (b - a);
}

// out[c] += w x component c at T, c < n (gridAddAt, :358-362): T clamped to the grid
inline void grid_add(const double *g, double T, int stride, int n, double w, double *out) {
  const double x = std::min(std::max(T - TGRID_MIN, 0.0), (double)(TGRID_N - 1));
  const int i = std::min((int)std::floor(x), TGRID_N - 2);
  const double t = x - i;
  const double *a = g + (int64_t)i * stride, *b = a + stride;
  for (int c = 0; c < n; c++) out[c] += w * (a[c] + t * (b[c] - a[c]));
}

// the pieces of [min(Ta, Tb), max(Ta, Tb)] between grid points, in increasing
// T: fn(midpoint, start, end) (forGridSegments, :370-377)
template <class Fn>
inline void for_pieces(double Ta, double Tb, Fn fn) {
  const double lo = std::min(Ta, Tb), hi = std::max(Ta, Tb);
  for (double t0 = lo; t0 < hi - 1e-12;) {
    const double t1 = std::min(hi, TGRID_MIN + std::floor(t0 - TGRID_MIN + 1e-9) + 1);
    fn(0.5 * (t0 + t1), t0, t1);
    t0 = t1;
  }
}

// mean of components 0..n-1 over the step (gridMean, :380-387); a zero-length
// step takes the value at Ta
inline void step_mean(const double *g, int stride, int n, double Ta, double Tb, double *out) {
  for (int c = 0; c < n; c++) out[c] = 0;
  const double L = std::fabs(Tb - Ta);
  if (L < 1e-12) {
    grid_add(g, Ta, stride, n, 1, out);
    return;
  }
  for_pieces(Ta, Tb, [&](double tm, double t0, double t1) { grid_add(g, tm, stride, n, (t1 - t0) / L, out); });
}

// integral from Ta to Tb of D(T) dF(T) (gridStressIncrement, :394-409): D a
// stiffness grid Dg (36 per point) or, when Dg is null, the constant Dc; F the
// normal thermal strains (3 per point)
inline void step_stress(const double *Dg, const double *Dc, const double *F, double Ta, double Tb, double out[6]) {
  for (int r = 0; r < 6; r++) out[r] = 0;
  if (!Dg) {
    const double f0 = grid_value(F, Tb, 3, 0) - grid_value(F, Ta, 3, 0);
    const double f1 = grid_value(F, Tb, 3, 1) - grid_value(F, Ta, 3, 1);
    const double f2 = grid_value(F, Tb, 3, 2) - grid_value(F, Ta, 3, 2);
    for (int r = 0; r < 6; r++) out[r] = Dc[6 * r] * f0 + Dc[6 * r + 1] * f1 + Dc[6 * r + 2] * f2;
    return;
  }
  const double sg = Tb >= Ta ? 1 : -1;
  for_pieces(Ta, Tb, [&](double tm, double t0, double t1) {
    double D[36];
    for (int q = 0; q < 36; q++) D[q] = 0;
    grid_add(Dg, tm, 36, 36, 1, D);
    const double f0 = grid_value(F, t1, 3, 0) - grid_value(F, t0, 3, 0);
    const double f1 = grid_value(F, t1, 3, 1) - grid_value(F, t0, 3, 1);
    const double f2 = grid_value(F, t1, 3, 2) - grid_value(F, t0, 3, 2);
    for (int r = 0; r < 6; r++) out[r] += sg * (D[6 * r] * f0 + D[6 * r + 1] * f1 + D[6 * r + 2] * f2);
  });
}

// thermal stress weights of a layered section over the step (sectionVstep,
// src/core/04-homogenize.js:212-221): V[j][r] = sum over sub-layers s of
// (integral of D_s dF_s)[r] x Ls[s][j]
inline void section_stress_weights(int ns, const double *D, const double *F, const double *Ls, double Ta, double Tb,
                                   double out[18]) {
  for (int q = 0; q < 18; q++) out[q] = 0;
  for (int s = 0; s < ns; s++) {
    double s0[6];
    step_stress(D + (int64_t)s * TGRID_N * 36, nullptr, F + (int64_t)s * TGRID_N * 3, Ta, Tb, s0);
    for (int r = 0; r < 6; r++)
      for (int j = 0; j < 3; j++) out[j * 6 + r] += s0[r] * Ls[3 * s + j];
  }
}

}  // namespace sd

// Fine operator on the matrix cores: y = A x (MODE 0) or y = b - A x (MODE 1)
// over every grid node, A the Jacobi-scaled, constraint-reduced operator
//   y_i = w_i sum_e sc_e (K_t(e) (w x)_e)_i   on free DOFs,  y_i = x_i on the others
// (MODE 1 is the multigrid cycle's residual, whose iterate x already holds w x:
// the staging takes x as it is and constrained rows give b_i)
// (the element-by-element form of the assembled product of
// src/core/09-assembly.js:96-108, 151-165, 445-464, with the condensed element
// matrices of src/core/09-assembly.js:368-424 shared per element type).
//
// Work split.  The grid's node columns are cut into tiles of TT x TT and the
// node planes into chunks of ZC; a workgroup owns one (tile, chunk).  It
// stages w*x of the tile's (TT+2)^2 node columns over the chunk's planes and
// one plane either side in LDS, then multiplies every element touching an
// owned node (the (TT+1)^2 cell columns around the tile, over the chunk's
// layers and the layer below; cells on a tile's rim or below a chunk are also
// computed by the neighbouring workgroup) by its type's 24 x 24 table on the
// matrix cores: the elements are listed by type (built once per mesh, below)
// in groups of 16, and one group is the product K_t [24 x 24] x X [24 x 16]
// = two 16-row blocks x six 4-deep steps of v_mfma_{f32,f64}_16x16x4 (rows
// 24..31 of the second block are zero).  The 24 element rows land in LDS and
// each owned node row sums its cells' rows in a fixed order (cells of the
// layer below, then of its own layer; slot-0 elements, then slot-1), so the
// result does not depend on the schedule.
//
// The same lists and index rules run on the host (elem_op_host) in the
// host-emulation build, with every LDS index checked.
#pragma once
#include <limits>

namespace mop {

constexpr int TT = 4;                      // owned node columns per tile side
constexpr int ZC = 8;                      // owned node planes per chunk
constexpr int NCC = TT + 1;                // cell columns per tile side
constexpr int NNC = TT + 2;                // staged node columns per tile side
constexpr int ZL = ZC + 1;                 // cell layers of a chunk (its planes' layers and the one below)
constexpr int ZN = ZC + 2;                 // staged node planes of a chunk
constexpr int GRP = 16;                    // elements per group (the product's N)
constexpr int CELLS = NCC * NCC * ZL;      // local cells of a chunk
constexpr int XSN = NNC * NNC * ZN;        // staged nodes of a chunk
constexpr int ROWS_PER_COL = ZC * 3;       // node rows of one owned column in a chunk
constexpr int ITEMS = TT * TT * ROWS_PER_COL;
constexpr int THREADS = 256;
static_assert(ITEMS <= 2 * THREADS, "two node rows per thread");

// chunk z owns node planes k0 = z * ZC .. k0 + ZC - 1; its local cell layer
// dl is grid layer k0 - 1 + dl, its local node plane dn is grid plane k0 - 1 + dn
FE_HD int lcell(int u, int v, int dl) { return (u * NCC + v) * ZL + dl; }
FE_HD int xnode(int a, int b, int dn) { return (a * NNC + b) * ZN + dn; }

// Element groups of every (tile, chunk, slot): elements sorted by type (stable
// in local cell order), cut into groups of GRP, the last group of a type
// padded with -1.  off[(tile * nchunk + chunk) * 2 + slot] is the first group
// and mask[((tile * nchunk + chunk) * 2 + slot) * MASKW ...] marks the listed cells.
constexpr int MASKW = (CELLS + 31) / 32;  // words of a chunk's cell mask

struct Lists {
  int ntx = 0, nty = 0, nchunk = 0;
  std::vector<int32_t> type;
  std::vector<int16_t> cell;
  std::vector<int32_t> off;
  std::vector<uint32_t> mask;  // MASKW per (tile, chunk, slot): the local cells that hold an element
};

inline void build_lists(const fe::Grid &g, const std::vector<int32_t> &cellType, Lists &L) {
  L.ntx = (g.nx + TT - 1) / TT;
  L.nty = (g.ny + TT - 1) / TT;
  L.nchunk = (g.nz + ZC - 1) / ZC;
  L.type.clear();
  L.cell.clear();
  L.off.assign((size_t)L.ntx * L.nty * L.nchunk * 2 + 1, 0);
  L.mask.assign((size_t)L.ntx * L.nty * L.nchunk * 2 * MASKW, 0u);
  std::vector<std::pair<int32_t, int16_t>> el;
  size_t slot = 0;
  for (int tile = 0; tile < L.ntx * L.nty; tile++) {
    const int i0 = (tile / L.nty) * TT, j0 = (tile % L.nty) * TT;
    for (int z = 0; z < L.nchunk; z++)
      for (int ph = 0; ph < 2; ph++) {
        L.off[slot++] = (int32_t)L.type.size();
        el.clear();
        for (int u = 0; u < NCC; u++)
          for (int v = 0; v < NCC; v++)
            for (int dl = 0; dl < ZL; dl++) {
              const int ii = i0 - 1 + u, jj = j0 - 1 + v, kk = z * ZC - 1 + dl;
              if (ii < 0 || ii >= g.ncx || jj < 0 || jj >= g.ncy || kk < 0 || kk >= g.ncz) continue;
              const int64_t c = ((int64_t)ii * g.ncy + jj) * g.ncz + kk;
              const int32_t t = cellType[2 * c + ph];
              if (t < 0) continue;
              const int lc = lcell(u, v, dl);
              el.push_back({t, (int16_t)lc});
              L.mask[(slot - 1) * MASKW + lc / 32] |= 1u << (lc % 32);
            }
        std::stable_sort(el.begin(), el.end(), [](const std::pair<int32_t, int16_t> &a,
                                                  const std::pair<int32_t, int16_t> &b) { return a.first < b.first; });
        for (size_t a = 0; a < el.size();) {
          size_t b = a;
          while (b < el.size() && el[b].first == el[a].first) b++;
          for (size_t s = a; s < b; s += GRP) {
            L.type.push_back(el[a].first);
            for (int q = 0; q < GRP; q++) L.cell.push_back(s + q < b ? el[s + q].second : (int16_t)-1);
          }
          a = b;
        }
      }
  }
  L.off[slot] = (int32_t)L.type.size();
  if (L.type.size() > (size_t)INT32_MAX / GRP) throw be::Error("too many element groups");
}

// node row `it` of a chunk: owned column nc (tile-local a = nc / TT + 1,
// b = nc % TT + 1 in staged coordinates), local plane dn = 1 + it's plane, component p
FE_HD void item_of(int it, int &nc, int &dn, int &p) {
  nc = it / ROWS_PER_COL;
  const int rem = it % ROWS_PER_COL;
  dn = 1 + rem / 3;
  p = rem % 3;
}

// sum of the cells' rows at staged node (a, b, dn), component p, in the fixed
// order (layer below, then own layer; within a layer by cell column), over
// the cells the mask marks (the others, outside the grid or empty, hold none)
template <typename T>
FE_HD T node_sum(const T *Y, const uint32_t *mask, int a, int b, int dn, int p) {
  T s = 0;
  for (int lk = 1; lk >= 0; lk--) {
    const int dl = dn - lk;
    for (int du = 0; du < 2; du++)
      for (int dv = 0; dv < 2; dv++) {
        const int corner = fe::corner_of(1 - du, 1 - dv, lk);
        const int lc = lcell(a - 1 + du, b - 1 + dv, dl);
        if (mask[lc >> 5] >> (lc & 31) & 1u) s += Y[lc * 24 + 3 * corner + p];
      }
  }
  return s;
}

#if FCBGA_GPU
template <typename T>
struct Mfma;
template <>
struct Mfma<float> {
  typedef float V4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ V4 op(float a, float b, V4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // row of the 16 x 16 result held in register v of `lane` (its column is lane % 16)
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};
template <>
struct Mfma<double> {
  typedef double V4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ V4 op(double a, double b, V4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};

// The element products of one slot's groups, wave by wave: each wave takes a
// contiguous run of the groups (one table load per type run), the next
// group's list entries are loaded while the current one multiplies.
template <typename T>
struct GroupRun {
  int gi, ge, tcur, tn, en;
  T av0[6], av1[6];
  __device__ __forceinline__ void start(int g0, int g1, int wave, const int32_t *__restrict__ gtype,
                                        const int16_t *__restrict__ gcell, const T *__restrict__ Kc, int jel,
                                        int kq) {
    const int n = g1 - g0;
    gi = g0 + (n * wave) / (THREADS / 64);
    ge = g0 + (n * (wave + 1)) / (THREADS / 64);
    tcur = -1;
    if (gi < ge) {
      tn = gtype[gi];
      en = gcell[(int64_t)gi * GRP + jel];
      load_table(tn, Kc, jel, kq);
    }
  }
  __device__ __forceinline__ void load_table(int t, const T *__restrict__ Kc, int jel, int kq) {
    const T *K = Kc + (int64_t)t * 576;
#pragma unroll
    for (int s = 0; s < 6; s++) {
      const int col = 4 * s + kq;
      av0[s] = K[jel * 24 + col];
      av1[s] = jel < 8 ? K[(16 + jel) * 24 + col] : T(0);
    }
    tcur = t;
  }
  __device__ __forceinline__ void run(const fe::Grid &g, int i0, int j0, int k0, int ph,
                                      const int32_t *__restrict__ gtype, const int16_t *__restrict__ gcell,
                                      const T *__restrict__ cs, const T *__restrict__ Kc, const T *xs, T *Y,
                                      int lane, int jel, int kq) {
    for (; gi < ge; gi++) {
      const int t = tn, e = en;
      if (gi + 1 < ge) {
        tn = gtype[gi + 1];
        en = gcell[(int64_t)(gi + 1) * GRP + jel];
      }
      if (t != tcur) load_table(t, Kc, jel, kq);
      const int ee = e >= 0 ? e : 0;
      const int dl = ee % ZL, uv = ee / ZL, u = uv / NCC, v = uv % NCC;
      T sc = 0;
      if (e >= 0) {
        const int64_t c = ((int64_t)(i0 - 1 + u) * g.ncy + (j0 - 1 + v)) * g.ncz + k0 - 1 + dl;
        sc = cs[2 * c + ph];
      }
      typename Mfma<T>::V4 c0 = {0, 0, 0, 0}, c1 = {0, 0, 0, 0};
#pragma unroll
      for (int s = 0; s < 6; s++) {
        const int col = 4 * s + kq, cb = col / 3, p = col % 3;
        const T xb = xs[xnode(u + fe::corner_li(cb), v + fe::corner_lj(cb), dl + fe::corner_lk(cb)) * 3 + p];
        const T bv = e >= 0 ? xb : T(0);
        c0 = Mfma<T>::op(av0[s], bv, c0);
        c1 = Mfma<T>::op(av1[s], bv, c1);
      }
      if (e >= 0) {
        T *Ye = Y + e * 24;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          Ye[Mfma<T>::row(lane, r)] = sc * c0[r];
          const int r1 = 16 + Mfma<T>::row(lane, r);
          if (r1 < 24) Ye[r1] = sc * c1[r];
        }
      }
    }
  }
};

template <typename T, int MODE>
__global__ void __launch_bounds__(THREADS) k_elem_mfma(fe::Grid g, int ntx, int nty, int nchunk,
                                                       const int32_t *__restrict__ gtype,
                                                       const int16_t *__restrict__ gcell,
                                                       const int32_t *__restrict__ off,
                                                       const uint32_t *__restrict__ gmask, const T *__restrict__ cs,
                                                       const T *__restrict__ Kc, const T *__restrict__ w,
                                                       const T *__restrict__ x, const T *__restrict__ bvec,
                                                       T *__restrict__ y) {
  __shared__ T xs[XSN * 3];
  __shared__ T Y[CELLS * 24];
  __shared__ uint32_t msk[2 * MASKW];
  const int q0 = xcd_block(blockIdx.x, gridDim.x);
  if (q0 >= ntx * nty * nchunk) return;
  const int tile = q0 / nchunk, z = q0 % nchunk;
  const int i0 = (tile / nty) * TT, j0 = (tile % nty) * TT, k0 = z * ZC;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int jel = lane & 15, kq = lane >> 4;
  const int g0 = off[2 * q0], g1 = off[2 * q0 + 1], g2 = off[2 * q0 + 2];
  // every global load of the block first: slot-0 group entries and table,
  // the cell masks, the staged w*x, the owned rows' w, x and b
  GroupRun<T> run;
  run.start(g0, g1, wave, gtype, gcell, Kc, jel, kq);
  if (tid < 2 * MASKW) msk[tid] = gmask[(int64_t)2 * q0 * MASKW + tid];
  int64_t gno[2];
  T wo[2], xo[2], bo[2];
  for (int h = 0; h < 2; h++) {
    const int it = tid + h * THREADS;
    int nc, dn, p;
    item_of(it < ITEMS ? it : 0, nc, dn, p);
    const int i = i0 + nc / TT, j = j0 + nc % TT, k = k0 - 1 + dn;
    const bool ok = it < ITEMS && i < g.nx && j < g.ny && k < g.nz;
    gno[h] = ok ? (((int64_t)i * g.ny + j) * g.nz + k) * 3 + p : -1;
    wo[h] = ok ? w[gno[h]] : T(0);
    xo[h] = ok ? x[gno[h]] : T(0);
    bo[h] = ok && MODE ? bvec[gno[h]] : T(0);
  }
  for (int q = tid; q < XSN; q += THREADS) {
    const int dn = q % ZN, ab = q / ZN, a = ab / NNC, b = ab % NNC;
    const int i = i0 - 1 + a, j = j0 - 1 + b, k = k0 - 1 + dn;
    const bool in = i >= 0 && i < g.nx && j >= 0 && j < g.ny && k >= 0 && k < g.nz;
    const int64_t gn = in ? (((int64_t)i * g.ny + j) * g.nz + k) * 3 : 0;
    for (int p = 0; p < 3; p++) xs[q * 3 + p] = in ? (MODE ? x[gn + p] : w[gn + p] * x[gn + p]) : T(0);
  }
  __syncthreads();
  T acc[2] = {T(0), T(0)};
  for (int ph = 0; ph < 2; ph++) {
    if (ph == 1) {
      if (g1 == g2) break;  // no slot-1 elements (the same for the whole block)
      __syncthreads();      // the slot-0 sums have read Y
      run.start(g1, g2, wave, gtype, gcell, Kc, jel, kq);
    }
    run.run(g, i0, j0, k0, ph, gtype, gcell, cs, Kc, xs, Y, lane, jel, kq);
    __syncthreads();
    for (int h = 0; h < 2; h++) {
      const int it = tid + h * THREADS;
      if (it >= ITEMS) continue;
      int nc, dn, p;
      item_of(it, nc, dn, p);
      acc[h] += node_sum<T>(Y, msk + ph * MASKW, nc / TT + 1, nc % TT + 1, dn, p);
    }
  }
  for (int h = 0; h < 2; h++) {
    if (gno[h] < 0) continue;
    const T out = wo[h] != T(0) ? wo[h] * acc[h] : (MODE ? T(0) : xo[h]);
    y[gno[h]] = MODE ? bo[h] - out : out;
  }
}
#else
// Host form of the kernel above: the same lists, staging, element rows and
// node sums (products summed in column order), every LDS index checked.
template <typename T, int MODE>
void elem_op_host(const fe::Grid &g, int ntx, int nty, int nchunk, const int32_t *gtype, const int16_t *gcell,
                  const int32_t *off, const uint32_t *gmask, const T *cs, const T *Kc, const T *w, const T *x,
                  const T *bvec, T *y) {
  std::vector<T> xs(XSN * 3), Y(CELLS * 24), acc(ITEMS);
  const auto chk = [](int idx, int n) {
    if (idx < 0 || idx >= n) throw be::Error("matrix-core operator: local index out of range");
    return idx;
  };
  for (int q0 = 0; q0 < ntx * nty * nchunk; q0++) {
    const int tile = q0 / nchunk, z = q0 % nchunk;
    const int i0 = (tile / nty) * TT, j0 = (tile % nty) * TT, k0 = z * ZC;
    for (int q = 0; q < XSN; q++) {
      const int dn = q % ZN, ab = q / ZN, a = ab / NNC, b = ab % NNC;
      const int i = i0 - 1 + a, j = j0 - 1 + b, k = k0 - 1 + dn;
      const bool in = i >= 0 && i < g.nx && j >= 0 && j < g.ny && k >= 0 && k < g.nz;
      const int64_t gn = in ? (((int64_t)i * g.ny + j) * g.nz + k) * 3 : 0;
      for (int p = 0; p < 3; p++) xs[chk(q * 3 + p, XSN * 3)] = in ? (MODE ? x[gn + p] : w[gn + p] * x[gn + p]) : T(0);
    }
    std::fill(acc.begin(), acc.end(), T(0));
    for (int ph = 0; ph < 2; ph++) {
      const int g0 = off[q0 * 2 + ph], g1 = off[q0 * 2 + ph + 1];
      if (g0 == g1) continue;
      // cells the mask leaves out must never be read: fill them with NaN
      std::fill(Y.begin(), Y.end(), std::numeric_limits<T>::quiet_NaN());
      for (int gi = g0; gi < g1; gi++) {
        const T *K = Kc + (int64_t)gtype[gi] * 576;
        for (int jel = 0; jel < GRP; jel++) {
          const int e = gcell[(int64_t)gi * GRP + jel];
          if (e < 0) continue;
          chk(e, CELLS);
          if (!(gmask[(int64_t)(q0 * 2 + ph) * MASKW + e / 32] >> (e % 32) & 1u))
            throw be::Error("matrix-core operator: listed cell missing from the mask");
          const int dl = e % ZL, uv = e / ZL, u = uv / NCC, v = uv % NCC;
          const int ii = i0 - 1 + u, jj = j0 - 1 + v, kk = k0 - 1 + dl;
          if (ii < 0 || ii >= g.ncx || jj < 0 || jj >= g.ncy || kk < 0 || kk >= g.ncz)
            throw be::Error("matrix-core operator: listed cell outside the grid");
          const int64_t c = ((int64_t)ii * g.ncy + jj) * g.ncz + kk;
          const T sc = cs[2 * c + ph];
          T xe[24];
          for (int col = 0; col < 24; col++) {
            const int cb = col / 3, p = col % 3;
            xe[col] = xs[chk(xnode(u + fe::corner_li(cb), v + fe::corner_lj(cb), dl + fe::corner_lk(cb)) * 3 + p,
                             XSN * 3)];
          }
          for (int r = 0; r < 24; r++) {
            T s = 0;
            for (int col = 0; col < 24; col++) s += K[r * 24 + col] * xe[col];
            Y[chk(e * 24 + r, CELLS * 24)] = sc * s;
          }
        }
      }
      for (int it = 0; it < ITEMS; it++) {
        int nc, dn, p;
        item_of(it, nc, dn, p);
        const int a = nc / TT + 1, b = nc % TT + 1;
        chk(lcell(a - 1, b - 1, dn - 1), CELLS);
        chk(lcell(a, b, dn), CELLS);
        acc[it] += node_sum<T>(Y.data(), gmask + (int64_t)(q0 * 2 + ph) * MASKW, a, b, dn, p);
      }
    }
    for (int it = 0; it < ITEMS; it++) {
      int nc, dn, p;
      item_of(it, nc, dn, p);
      const int i = i0 + nc / TT, j = j0 + nc % TT, k = k0 - 1 + dn;
      if (i >= g.nx || j >= g.ny || k >= g.nz) continue;
      const int64_t gn = (((int64_t)i * g.ny + j) * g.nz + k) * 3 + p;
      const T wi = w[gn];
      const T out = wi != T(0) ? wi * acc[it] : (MODE ? T(0) : x[gn]);
      y[gn] = MODE ? bvec[gn] - out : out;
    }
  }
}
#endif

}  // namespace mop

// fpc_oracle.cpp -- CPU oracle for the per-timestep path (TEST INFRASTRUCTURE ONLY).
//
// A plain C++ restatement of the reference algorithm, sequential unless
// orc_set_threads(n > 1) spreads its solver and stage loops over OpenMP
// threads (bench.py's CPU baseline does; the tests run it sequentially).  Tests,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg use it as the checker;
// nothing in the product library links, loads or calls it.
//
// What it restates (reference files, JD63021/ibm-3d-fpc):
//   * host sparse type and algebra with std::unordered_map rows  common.hpp:136-228
//   * geometry fractions, masks, pruning                         immersed_geometry.hpp:1-160
//   * 3-D extrusion, 1-based id maps                             main.cu:33-85
//   * u/v/w diffusion, D, G, Ap = D*G, row scaling, sign flip    immersed_operators.hpp:1-287, 737-748; main.cu:96-103
//   * flat_of_id + AoS 7-point embedding                         immersed_operators.hpp:397-447; solver_helpers.hpp:51-84, 227-329
//   * CSR with sorted columns                                    gpu_dense_ops.hpp:32-48
//   * the timestep kernels, element by element                   gpu_dense_ops.hpp:1-106
//   * the time-step driver                                       main.cu:249-363
//   * HYPRE Struct GMRES (right-preconditioned, modified Gram-Schmidt, Givens,
//     true-residual restart check) and Struct Jacobi (one sweep from zero
//     gives x = b/diag) -- HYPRE is not vendored in the reference; the control
//     flow follows HYPRE's published krylov/gmres.c algorithm (version unpinned).
//   * a PFMG-style semicoarsening V-cycle: the exact PFMG internals (RAP type 1
//     formula, relaxation weights, skip-relax) are unpinned, so the oracle states
//     the build's own definition, which the GPU path implements independently.
//   * drag/lift and the seeded-noise initial condition: not in the reference;
//     defined here and in the product identically.
//
// Compile with -ffp-contract=off so host arithmetic rounds exactly as written.

#include "fpc_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include <omp.h>

namespace {

typedef unsigned char u8;

// ---------------------------------------------------------------- indexing
// Staggered full-array layouts, x fastest (common.hpp:230-238).
// Host threads of the solver loops (orc_set_threads; 1 by default).  Loops whose
// iterations write disjoint outputs run as OpenMP static partitions; reductions
// (dot) sum fixed per-thread partials in thread order, so results depend only on
// the thread count and with one thread are the sequential sums.
static int g_threads = 1;
#define ORC_PAR _Pragma("omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)")
#define ORC_PAR2 _Pragma("omp parallel for collapse(2) schedule(static) num_threads(g_threads) if (g_threads > 1)")

inline size_t ixP(int i, int j, int k, int nx, int ny) { return ((size_t)k * ny + j) * nx + i; }
inline size_t ixU(int i, int j, int k, int nx, int ny) { return ((size_t)k * ny + j) * (size_t)(nx + 1) + i; }
inline size_t ixV(int i, int j, int k, int nx, int ny) { return ((size_t)k * (ny + 1) + j) * nx + i; }
inline size_t ixW(int i, int j, int k, int nx, int ny) { return ((size_t)k * ny + j) * nx + i; }

inline double inlet_profile(double y, double z, double H, double Um) {
  // common.hpp:240-243
  return 16.0 * Um * y * z * (H - y) * (H - z) / std::pow(H, 4.0);
}

// Ray/segment-circle distance (common.hpp:262-280).
double seg_circle(double xP, double yP, double xS, double yS, double xc, double yc, double R) {
  double ddx = xS - xP, ddy = yS - yP;
  double a = ddx * ddx + ddy * ddy;
  if (a <= 0.0) return std::numeric_limits<double>::infinity();
  double ox = xP - xc, oy = yP - yc;
  double b = 2.0 * (ddx * ox + ddy * oy);
  double c = ox * ox + oy * oy - R * R;
  double disc = b * b - 4.0 * a * c;
  if (disc < 0.0) return std::numeric_limits<double>::infinity();
  double sd = std::sqrt(std::max(0.0, disc));
  double t1 = (-b - sd) / (2.0 * a), t2 = (-b + sd) / (2.0 * a);
  double t = std::numeric_limits<double>::infinity();
  if (t1 >= 0.0 && t1 <= 1.0) t = t1;
  else if (t2 >= 0.0 && t2 <= 1.0) t = t2;
  if (!std::isfinite(t)) return std::numeric_limits<double>::infinity();
  return std::sqrt(a) * t;
}

// ---------------------------------------------------------------- sparse type
// Rows are std::unordered_map<int,double> with 1-based columns (common.hpp:136-159).
// Keeping the real container keeps its iteration order, which fixes the
// summation order of Ap's diagonal and of the l1 row scales.
struct Sp {
  int nrows = 0, ncols = 0;
  std::vector<std::unordered_map<int, double>> rows;
  Sp() = default;
  Sp(int m, int n) : nrows(m), ncols(n), rows((size_t)m) {}
};

void sp_add(Sp &A, int r, int c, double v) {
  if (r < 1 || r > A.nrows || c < 1 || c > A.ncols) return;
  if (std::abs(v) < 1.0e-30) return;
  A.rows[(size_t)(r - 1)][c] += v;
}

// Rows are independent: the parallel loops below give each row the same
// insertion sequence as a serial run, hence the same container state.
void sp_compress(Sp &A, double tol = 1.0e-14) {
  ORC_PAR
  for (size_t i = 0; i < A.rows.size(); ++i) {
    auto &r = A.rows[i];
    for (auto it = r.begin(); it != r.end();) {
      if (std::abs(it->second) <= tol) it = r.erase(it);
      else ++it;
    }
  }
}

// Frees the row containers (millions of small hash maps) in parallel.
void sp_release(Sp &A) {
  ORC_PAR
  for (size_t i = 0; i < A.rows.size(); ++i) std::unordered_map<int, double>().swap(A.rows[i]);
  A.rows.clear();
  A.rows.shrink_to_fit();
}

Sp sp_mul(const Sp &A, const Sp &B) {  // common.hpp:175-189
  Sp C(A.nrows, B.ncols);
  ORC_PAR
  for (int i = 0; i < A.nrows; ++i)
    for (const auto &ak : A.rows[(size_t)i]) {
      int k = ak.first;
      if (k < 1 || k > B.nrows) continue;
      for (const auto &bk : B.rows[(size_t)(k - 1)]) sp_add(C, i + 1, bk.first, ak.second * bk.second);
    }
  sp_compress(C);
  return C;
}

void sp_row_scale(const Sp &A, Sp &As, std::vector<double> &scale) {  // immersed_operators.hpp:737-748
  scale.assign((size_t)A.nrows + 1, 1.0);
  As = Sp(A.nrows, A.ncols);
  ORC_PAR
  for (int i = 1; i <= A.nrows; ++i) {
    double s = 0.0;
    for (const auto &kv : A.rows[(size_t)(i - 1)]) s += std::abs(kv.second);
    if (s < 1.0e-14) s = 1.0;
    scale[(size_t)i] = 1.0 / s;
    for (const auto &kv : A.rows[(size_t)(i - 1)]) sp_add(As, i, kv.first, scale[(size_t)i] * kv.second);
  }
  sp_compress(As);
}

struct CSR {
  int nrows = 0;
  std::vector<int> rowptr, colind;
  std::vector<double> vals;
};

CSR sp_to_csr(const Sp &A) {  // gpu_dense_ops.hpp:32-48 (columns sorted ascending)
  CSR M;
  M.nrows = A.nrows;
  M.rowptr.assign((size_t)A.nrows + 1, 0);
  for (int i = 0; i < A.nrows; ++i) M.rowptr[(size_t)i + 1] = M.rowptr[(size_t)i] + (int)A.rows[(size_t)i].size();
  M.colind.assign((size_t)M.rowptr[(size_t)A.nrows], 0);
  M.vals.assign(M.colind.size(), 0.0);
  ORC_PAR
  for (int i = 0; i < A.nrows; ++i) {
    std::vector<std::pair<int, double>> row(A.rows[(size_t)i].begin(), A.rows[(size_t)i].end());
    std::sort(row.begin(), row.end(), [](const std::pair<int, double> &a, const std::pair<int, double> &b) {
      return a.first < b.first;
    });
    size_t o = (size_t)M.rowptr[(size_t)i];
    for (auto &kv : row) { M.colind[o] = kv.first; M.vals[o] = kv.second; ++o; }
  }
  return M;
}

void csr_spmv(const CSR &M, const double *x, double *y) {
  ORC_PAR
  for (int r = 0; r < M.nrows; ++r) {
    double s = 0.0;
    for (int jj = M.rowptr[r]; jj < M.rowptr[r + 1]; ++jj) s += M.vals[jj] * x[M.colind[jj]];
    y[r + 1] = s;
  }
}

// ---------------------------------------------------------------- setup
struct Setup {
  orc_params par{};
  int nx = 0, ny = 0, nz = 0;
  double dx = 0, dy = 0, dz = 0;
  std::vector<double> beta2d, alpha_u2d, alpha_v2d;
  std::vector<u8> fluid2d, active_u2d, active_v2d, active_qu2d, active_qv2d;
  int n_pruned = 0;
  std::vector<double> beta_p, alpha_u, alpha_v, alpha_w;
  std::vector<u8> fluid_p, active_u, active_v, active_w, active_qu, active_qv, active_qw;
  std::vector<int> id_u, id_v, id_w, id_p, id_qu, id_qv, id_qw;
  int Nu = 0, Nv = 0, Nw = 0, Np = 0, Nqu = 0, Nqv = 0, Nqw = 0;
  std::vector<double> rhs_u_bc, rhs_v_bc, rhs_w_bc, div_bc;
  std::vector<double> scale_u, scale_v, scale_w;
  std::vector<double> Au_vals, Av_vals, Aw_vals, Ap_vals;  // AoS, 7 per box point
  CSR D, G;
  std::vector<double> uin;
};

// immersed_geometry.hpp:1-27
void geom_beta(Setup &S) {
  const orc_params &p = S.par;
  int nx = S.nx, ny = S.ny, ns = p.nsub_beta;
  std::vector<double> sub((size_t)ns);
  for (int k = 0; k < ns; ++k) sub[(size_t)k] = ((double)k + 0.5) / (double)ns;
  S.beta2d.assign((size_t)nx * ny, 0.0);
  for (int i = 1; i <= nx; ++i) {
    double x0 = (i - 1) * S.dx;
    for (int j = 1; j <= ny; ++j) {
      double y0 = (j - 1) * S.dy;
      int cnt = 0;
      for (int a = 0; a < ns; ++a)
        for (int b = 0; b < ns; ++b) {
          double xs = x0 + sub[(size_t)a] * S.dx;
          double ys = y0 + sub[(size_t)b] * S.dy;
          double rsq = (xs - p.xcirc) * (xs - p.xcirc) + (ys - p.ycirc) * (ys - p.ycirc);
          if (rsq > p.Rcirc * p.Rcirc) cnt++;
        }
      S.beta2d[(size_t)(i - 1) + (size_t)(j - 1) * nx] = (double)cnt / (double)(ns * ns);
    }
  }
}

// immersed_geometry.hpp:29-73
void geom_alpha(Setup &S) {
  const orc_params &p = S.par;
  int nx = S.nx, ny = S.ny;
  double R = p.Rcirc;
  S.alpha_u2d.assign((size_t)(nx + 1) * ny, 1.0);
  for (int i = 1; i <= nx + 1; ++i) {
    double x = (i - 1) * S.dx;
    for (int j = 1; j <= ny; ++j) {
      double y0 = (j - 1) * S.dy, y1 = j * S.dy, val = 1.0;
      if (std::abs(x - p.xcirc) < R) {
        double h = std::sqrt(std::max(0.0, R * R - (x - p.xcirc) * (x - p.xcirc)));
        double lo = p.ycirc - h, hi = p.ycirc + h;
        double ov = std::max(0.0, std::min(y1, hi) - std::max(y0, lo));
        val = std::max(0.0, std::min(1.0, 1.0 - ov / S.dy));
      }
      S.alpha_u2d[(size_t)(i - 1) + (size_t)(j - 1) * (nx + 1)] = val;
    }
  }
  S.alpha_v2d.assign((size_t)nx * (ny + 1), 1.0);
  for (int j = 1; j <= ny + 1; ++j) {
    double y = (j - 1) * S.dy;
    for (int i = 1; i <= nx; ++i) {
      double x0 = (i - 1) * S.dx, x1 = i * S.dx, val = 1.0;
      if (std::abs(y - p.ycirc) < R) {
        double h = std::sqrt(std::max(0.0, R * R - (y - p.ycirc) * (y - p.ycirc)));
        double lo = p.xcirc - h, hi = p.xcirc + h;
        double ov = std::max(0.0, std::min(x1, hi) - std::max(x0, lo));
        val = std::max(0.0, std::min(1.0, 1.0 - ov / S.dx));
      }
      S.alpha_v2d[(size_t)(i - 1) + (size_t)(j - 1) * nx] = val;
    }
  }
}

// immersed_geometry.hpp:75-115
void geom_masks(Setup &S) {
  const orc_params &p = S.par;
  int nx = S.nx, ny = S.ny;
  double R2 = p.Rcirc * p.Rcirc;
  auto fl = [&](int i1, int j1) { return S.fluid2d[(size_t)(i1 - 1) + (size_t)(j1 - 1) * nx] != 0; };
  S.active_u2d.assign((size_t)(nx + 1) * ny, 0);
  S.active_qu2d.assign((size_t)(nx + 1) * ny, 0);
  S.active_v2d.assign((size_t)nx * (ny + 1), 0);
  S.active_qv2d.assign((size_t)nx * (ny + 1), 0);
  for (int i = 2; i <= nx; ++i) {
    double x = (i - 1) * S.dx;
    for (int j = 1; j <= ny; ++j) {
      double y = (j - 0.5) * S.dy;
      bool out = ((x - p.xcirc) * (x - p.xcirc) + (y - p.ycirc) * (y - p.ycirc)) > R2;
      bool adj = fl(i - 1, j) || fl(i, j);
      size_t f = (size_t)(i - 1) + (size_t)(j - 1) * (nx + 1);
      S.active_u2d[f] = (u8)(out && adj);
      S.active_qu2d[f] = (u8)(out && adj && S.alpha_u2d[f] > p.alpha_tol);
    }
  }
  for (int j = 1; j <= ny; ++j) S.active_qu2d[(size_t)nx + (size_t)(j - 1) * (nx + 1)] = (u8)fl(nx, j);
  for (int i = 1; i <= nx; ++i) {
    double x = (i - 0.5) * S.dx;
    for (int j = 2; j <= ny; ++j) {
      double y = (j - 1) * S.dy;
      bool out = ((x - p.xcirc) * (x - p.xcirc) + (y - p.ycirc) * (y - p.ycirc)) > R2;
      bool adj = fl(i, j - 1) || fl(i, j);
      size_t f = (size_t)(i - 1) + (size_t)(j - 1) * nx;
      S.active_v2d[f] = (u8)(out && adj);
      S.active_qv2d[f] = (u8)(out && adj && S.alpha_v2d[f] > p.alpha_tol);
    }
  }
}

// immersed_geometry.hpp:117-160 (fix-point pruning of isolated columns)
void geom_prune(Setup &S) {
  int nx = S.nx, ny = S.ny;
  S.fluid2d.assign((size_t)nx * ny, 0);
  for (size_t c = 0; c < S.fluid2d.size(); ++c) S.fluid2d[c] = (u8)(S.beta2d[c] > S.par.beta_tol);
  S.n_pruned = 0;
  for (;;) {
    geom_masks(S);
    std::vector<u8> rm((size_t)nx * ny, 0);
    bool any = false;
    for (int i = 1; i <= nx; ++i)
      for (int j = 1; j <= ny; ++j) {
        size_t c = (size_t)(i - 1) + (size_t)(j - 1) * nx;
        if (!S.fluid2d[c]) continue;
        int n = 0;
        if (S.active_qu2d[(size_t)i + (size_t)(j - 1) * (nx + 1)]) n++;  // east face (outflow when i == nx)
        if (i > 1 && S.active_qu2d[(size_t)(i - 1) + (size_t)(j - 1) * (nx + 1)]) n++;
        if (j < ny && S.active_qv2d[(size_t)(i - 1) + (size_t)j * nx]) n++;
        if (j > 1 && S.active_qv2d[(size_t)(i - 1) + (size_t)(j - 1) * nx]) n++;
        if (n == 0) { rm[c] = 1; any = true; }
      }
    if (!any) break;
    for (int i = 1; i <= nx; ++i)
      for (int j = 1; j <= ny; ++j) {
        size_t c = (size_t)(i - 1) + (size_t)(j - 1) * nx;
        if (rm[c]) { S.fluid2d[c] = 0; S.n_pruned++; }
      }
  }
  geom_masks(S);
}

// main.cu:33-85
void extrude_and_ids(Setup &S) {
  int nx = S.nx, ny = S.ny, nz = S.nz;
  S.beta_p.assign((size_t)nx * ny * nz, 0.0);
  S.fluid_p.assign((size_t)nx * ny * nz, 0);
  S.alpha_u.assign((size_t)(nx + 1) * ny * nz, 0.0);
  S.active_u.assign(S.alpha_u.size(), 0);
  S.active_qu.assign(S.alpha_u.size(), 0);
  S.alpha_v.assign((size_t)nx * (ny + 1) * nz, 0.0);
  S.active_v.assign(S.alpha_v.size(), 0);
  S.active_qv.assign(S.alpha_v.size(), 0);
  S.alpha_w.assign((size_t)nx * ny * (nz + 1), 1.0);
  S.active_w.assign(S.alpha_w.size(), 0);
  S.active_qw.assign(S.alpha_w.size(), 0);
  for (int k = 0; k < nz; ++k) {
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        S.beta_p[ixP(i, j, k, nx, ny)] = S.beta2d[(size_t)i + (size_t)j * nx];
        S.fluid_p[ixP(i, j, k, nx, ny)] = S.fluid2d[(size_t)i + (size_t)j * nx];
      }
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        size_t s = (size_t)i + (size_t)j * (nx + 1), d = ixU(i, j, k, nx, ny);
        S.alpha_u[d] = S.alpha_u2d[s];
        S.active_u[d] = S.active_u2d[s];
        S.active_qu[d] = S.active_qu2d[s];
      }
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t s = (size_t)i + (size_t)j * nx, d = ixV(i, j, k, nx, ny);
        S.alpha_v[d] = S.alpha_v2d[s];
        S.active_v[d] = S.active_v2d[s];
        S.active_qv[d] = S.active_qv2d[s];
      }
  }
  for (int kw = 0; kw <= nz; ++kw)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        bool adj = (kw >= 1 && S.fluid_p[ixP(i, j, kw - 1, nx, ny)]) || (kw <= nz - 1 && S.fluid_p[ixP(i, j, kw, nx, ny)]);
        S.active_w[ixW(i, j, kw, nx, ny)] = (u8)adj;
        S.active_qw[ixW(i, j, kw, nx, ny)] = (u8)adj;
      }
  auto ids = [](const std::vector<u8> &m, std::vector<int> &id) {
    id.assign(m.size(), 0);
    int c = 0;
    for (size_t t = 0; t < m.size(); ++t)
      if (m[t]) id[t] = ++c;
    return c;
  };
  S.Nu = ids(S.active_u, S.id_u);
  S.Nv = ids(S.active_v, S.id_v);
  S.Nw = ids(S.active_w, S.id_w);
  S.Np = ids(S.fluid_p, S.id_p);
  S.Nqu = ids(S.active_qu, S.id_qu);
  S.Nqv = ids(S.active_qv, S.id_qv);
  S.Nqw = ids(S.active_qw, S.id_qw);
}

// immersed_operators.hpp:1-62
void build_u(const Setup &S, Sp &A, std::vector<double> &bc) {
  const orc_params &p = S.par;
  int nx = S.nx, ny = S.ny, nz = S.nz;
  double dx = S.dx, dy = S.dy, dz = S.dz;
  double cx = p.nu / (dx * dx), cy = p.nu / (dy * dy), cz = p.nu / (dz * dz), at = p.rho / p.dt;
  A = Sp(S.Nu, S.Nu);
  bc.assign((size_t)S.Nu + 1, 0.0);
  auto act = [&](int i0, int j0, int k0) { return S.active_u[ixU(i0, j0, k0, nx, ny)] != 0; };
  auto idf = [&](int i0, int j0, int k0) { return S.id_u[ixU(i0, j0, k0, nx, ny)]; };
  auto seg = [&](double a, double b, double c, double d) { return seg_circle(a, b, c, d, p.xcirc, p.ycirc, p.Rcirc); };
  for (int i = 2; i <= nx; ++i) {
    double xP = (i - 1) * dx;
    for (int j = 1; j <= ny; ++j) {
      double yP = (j - 0.5) * dy;
      for (int k = 1; k <= nz; ++k) {
        if (!act(i - 1, j - 1, k - 1)) continue;
        int P = idf(i - 1, j - 1, k - 1);
        double zP = (k - 0.5) * dz, dg = at;
        if (i > 2 && act(i - 2, j - 1, k - 1)) { sp_add(A, P, idf(i - 2, j - 1, k - 1), -cx); dg += cx; }
        else if (i == 2) { dg += cx; bc[(size_t)P] += cx * inlet_profile(yP, zP, p.Ly, 1.0); }
        else dg += cx * (dx / seg(xP, yP, xP - dx, yP));
        if (i < nx && act(i, j - 1, k - 1)) { sp_add(A, P, idf(i, j - 1, k - 1), -cx); dg += cx; }
        else if (i != nx) dg += cx * (dx / seg(xP, yP, xP + dx, yP));
        if (j > 1 && act(i - 1, j - 2, k - 1)) { sp_add(A, P, idf(i - 1, j - 2, k - 1), -cy); dg += cy; }
        else dg += cy * ((j == 1) ? dy / (0.5 * dy) : dy / seg(xP, yP, xP, yP - dy));
        if (j < ny && act(i - 1, j, k - 1)) { sp_add(A, P, idf(i - 1, j, k - 1), -cy); dg += cy; }
        else dg += cy * ((j == ny) ? dy / (0.5 * dy) : dy / seg(xP, yP, xP, yP + dy));
        if (k > 1 && act(i - 1, j - 1, k - 2)) { sp_add(A, P, idf(i - 1, j - 1, k - 2), -cz); dg += cz; }
        else dg += 2.0 * cz;
        if (k < nz && act(i - 1, j - 1, k)) { sp_add(A, P, idf(i - 1, j - 1, k), -cz); dg += cz; }
        else dg += 2.0 * cz;
        sp_add(A, P, P, dg);
      }
    }
  }
  sp_compress(A);
}

// immersed_operators.hpp:64-121
void build_v(const Setup &S, Sp &A, std::vector<double> &bc) {
  const orc_params &p = S.par;
  int nx = S.nx, ny = S.ny, nz = S.nz;
  double dx = S.dx, dy = S.dy, dz = S.dz;
  double cx = p.nu / (dx * dx), cy = p.nu / (dy * dy), cz = p.nu / (dz * dz), at = p.rho / p.dt;
  A = Sp(S.Nv, S.Nv);
  bc.assign((size_t)S.Nv + 1, 0.0);
  auto act = [&](int i0, int j0, int k0) { return S.active_v[ixV(i0, j0, k0, nx, ny)] != 0; };
  auto idf = [&](int i0, int j0, int k0) { return S.id_v[ixV(i0, j0, k0, nx, ny)]; };
  auto seg = [&](double a, double b, double c, double d) { return seg_circle(a, b, c, d, p.xcirc, p.ycirc, p.Rcirc); };
  for (int i = 1; i <= nx; ++i) {
    double xP = (i - 0.5) * dx;
    for (int j = 2; j <= ny; ++j) {
      double yP = (j - 1) * dy;
      for (int k = 1; k <= nz; ++k) {
        if (!act(i - 1, j - 1, k - 1)) continue;
        int P = idf(i - 1, j - 1, k - 1);
        double dg = at;
        if (i > 1 && act(i - 2, j - 1, k - 1)) { sp_add(A, P, idf(i - 2, j - 1, k - 1), -cx); dg += cx; }
        else dg += cx * ((i == 1) ? dx / (0.5 * dx) : dx / seg(xP, yP, xP - dx, yP));
        if (i < nx && act(i, j - 1, k - 1)) { sp_add(A, P, idf(i, j - 1, k - 1), -cx); dg += cx; }
        else if (i != nx) dg += cx * (dx / seg(xP, yP, xP + dx, yP));
        if (j > 2 && act(i - 1, j - 2, k - 1)) { sp_add(A, P, idf(i - 1, j - 2, k - 1), -cy); dg += cy; }
        else dg += cy * ((j == 2) ? dy / dy : dy / seg(xP, yP, xP, yP - dy));
        if (j < ny && act(i - 1, j, k - 1)) { sp_add(A, P, idf(i - 1, j, k - 1), -cy); dg += cy; }
        else dg += cy * ((j == ny) ? dy / dy : dy / seg(xP, yP, xP, yP + dy));
        if (k > 1 && act(i - 1, j - 1, k - 2)) { sp_add(A, P, idf(i - 1, j - 1, k - 2), -cz); dg += cz; }
        else dg += 2.0 * cz;
        if (k < nz && act(i - 1, j - 1, k)) { sp_add(A, P, idf(i - 1, j - 1, k), -cz); dg += cz; }
        else dg += 2.0 * cz;
        sp_add(A, P, P, dg);
      }
    }
  }
  sp_compress(A);
}

// immersed_operators.hpp:123-177
void build_w(const Setup &S, Sp &A, std::vector<double> &bc) {
  const orc_params &p = S.par;
  int nx = S.nx, ny = S.ny, nz = S.nz;
  double dx = S.dx, dy = S.dy, dz = S.dz;
  double cx = p.nu / (dx * dx), cy = p.nu / (dy * dy), cz = p.nu / (dz * dz), at = p.rho / p.dt;
  A = Sp(S.Nw, S.Nw);
  bc.assign((size_t)S.Nw + 1, 0.0);
  auto act = [&](int i0, int j0, int k0) { return S.active_w[ixW(i0, j0, k0, nx, ny)] != 0; };
  auto idf = [&](int i0, int j0, int k0) { return S.id_w[ixW(i0, j0, k0, nx, ny)]; };
  auto seg = [&](double a, double b, double c, double d) { return seg_circle(a, b, c, d, p.xcirc, p.ycirc, p.Rcirc); };
  for (int i = 1; i <= nx; ++i) {
    double xP = (i - 0.5) * dx;
    for (int j = 1; j <= ny; ++j) {
      double yP = (j - 0.5) * dy;
      for (int k = 2; k <= nz; ++k) {
        if (!act(i - 1, j - 1, k - 1)) continue;
        int P = idf(i - 1, j - 1, k - 1);
        double dg = at;
        if (i > 1 && act(i - 2, j - 1, k - 1)) { sp_add(A, P, idf(i - 2, j - 1, k - 1), -cx); dg += cx; }
        else dg += cx * ((i == 1) ? dx / (0.5 * dx) : dx / seg(xP, yP, xP - dx, yP));
        if (i < nx && act(i, j - 1, k - 1)) { sp_add(A, P, idf(i, j - 1, k - 1), -cx); dg += cx; }
        else if (i != nx) dg += cx * (dx / seg(xP, yP, xP + dx, yP));
        if (j > 1 && act(i - 1, j - 2, k - 1)) { sp_add(A, P, idf(i - 1, j - 2, k - 1), -cy); dg += cy; }
        else dg += cy * ((j == 1) ? dy / (0.5 * dy) : dy / seg(xP, yP, xP, yP - dy));
        if (j < ny && act(i - 1, j, k - 1)) { sp_add(A, P, idf(i - 1, j, k - 1), -cy); dg += cy; }
        else dg += cy * ((j == ny) ? dy / (0.5 * dy) : dy / seg(xP, yP, xP, yP + dy));
        if (k > 2 && act(i - 1, j - 1, k - 2)) { sp_add(A, P, idf(i - 1, j - 1, k - 2), -cz); dg += cz; }
        else dg += cz;
        if (k < nz && act(i - 1, j - 1, k)) { sp_add(A, P, idf(i - 1, j - 1, k), -cz); dg += cz; }
        else dg += cz;
        sp_add(A, P, P, dg);
      }
    }
  }
  sp_compress(A);
}

// immersed_operators.hpp:179-287
void build_proj(const Setup &S, Sp &D, Sp &G, std::vector<double> &div_bc) {  // Ap = D G follows in setup_create
  const orc_params &p = S.par;
  int nx = S.nx, ny = S.ny, nz = S.nz;
  double dx = S.dx, dy = S.dy, dz = S.dz;
  int nQu = S.Nqu, nQv = S.Nqv, nF = S.Nqu + S.Nqv + S.Nqw;
  D = Sp(S.Np, nF);
  G = Sp(nF, S.Np);
  div_bc.assign((size_t)S.Np + 1, 0.0);
  for (int i = 1; i <= nx; ++i)
    for (int j = 1; j <= ny; ++j)
      for (int k = 1; k <= nz; ++k) {
        size_t c = ixP(i - 1, j - 1, k - 1, nx, ny);
        if (!S.fluid_p[c]) continue;
        int row = S.id_p[c];
        double beta = S.beta_p[c];
        size_t qe = ixU(i, j - 1, k - 1, nx, ny);  // east face (the outflow face when i == nx)
        if (S.active_qu[qe]) sp_add(D, row, S.id_qu[qe], S.alpha_u[qe] / (beta * dx));
        if (i > 1) {
          size_t qw = ixU(i - 1, j - 1, k - 1, nx, ny);
          if (S.active_qu[qw]) sp_add(D, row, S.id_qu[qw], -S.alpha_u[qw] / (beta * dx));
        } else {
          double y = (j - 0.5) * dy, z = (k - 0.5) * dz;
          div_bc[(size_t)row] -= inlet_profile(y, z, p.Ly, p.Um) / (beta * dx);
        }
        if (j < ny) {
          size_t qn = ixV(i - 1, j, k - 1, nx, ny);
          if (S.active_qv[qn]) sp_add(D, row, nQu + S.id_qv[qn], S.alpha_v[qn] / (beta * dy));
        }
        if (j > 1) {
          size_t qs = ixV(i - 1, j - 1, k - 1, nx, ny);
          if (S.active_qv[qs]) sp_add(D, row, nQu + S.id_qv[qs], -S.alpha_v[qs] / (beta * dy));
        }
        if (k < nz) {
          size_t qt = ixW(i - 1, j - 1, k, nx, ny);
          if (S.active_qw[qt]) sp_add(D, row, nQu + nQv + S.id_qw[qt], S.alpha_w[qt] / (beta * dz));
        }
        if (k > 1) {
          size_t qb = ixW(i - 1, j - 1, k - 1, nx, ny);
          if (S.active_qw[qb]) sp_add(D, row, nQu + nQv + S.id_qw[qb], -S.alpha_w[qb] / (beta * dz));
        }
      }
  for (int i = 2; i <= nx; ++i)
    for (int j = 1; j <= ny; ++j)
      for (int k = 1; k <= nz; ++k) {
        size_t q = ixU(i - 1, j - 1, k - 1, nx, ny);
        if (!S.active_qu[q]) continue;
        int row = S.id_qu[q];
        sp_add(G, row, S.id_p[ixP(i - 2, j - 1, k - 1, nx, ny)], -1.0 / dx);
        sp_add(G, row, S.id_p[ixP(i - 1, j - 1, k - 1, nx, ny)], +1.0 / dx);
      }
  for (int j = 1; j <= ny; ++j)
    for (int k = 1; k <= nz; ++k) {
      size_t q = ixU(nx, j - 1, k - 1, nx, ny);
      if (!S.active_qu[q]) continue;
      sp_add(G, S.id_qu[q], S.id_p[ixP(nx - 1, j - 1, k - 1, nx, ny)], -2.0 / dx);
    }
  for (int i = 1; i <= nx; ++i)
    for (int j = 2; j <= ny; ++j)
      for (int k = 1; k <= nz; ++k) {
        size_t q = ixV(i - 1, j - 1, k - 1, nx, ny);
        if (!S.active_qv[q]) continue;
        int row = nQu + S.id_qv[q];
        sp_add(G, row, S.id_p[ixP(i - 1, j - 2, k - 1, nx, ny)], -1.0 / dy);
        sp_add(G, row, S.id_p[ixP(i - 1, j - 1, k - 1, nx, ny)], +1.0 / dy);
      }
  for (int i = 1; i <= nx; ++i)
    for (int j = 1; j <= ny; ++j)
      for (int k = 2; k <= nz; ++k) {
        size_t q = ixW(i - 1, j - 1, k - 1, nx, ny);
        if (!S.active_qw[q]) continue;
        int row = nQu + nQv + S.id_qw[q];
        sp_add(G, row, S.id_p[ixP(i - 1, j - 1, k - 2, nx, ny)], -1.0 / dz);
        sp_add(G, row, S.id_p[ixP(i - 1, j - 1, k - 1, nx, ny)], +1.0 / dz);
      }
  sp_compress(D);
  sp_compress(G);
}

// Embedding of an active-row matrix into a 7-point AoS box (solver_helpers.hpp:51-84, 227-329).
// box(i,j,k) -> full flat index of the row's face/cell; 'full' dims give the decode.
void embed(const Sp &A, const std::vector<int> &id_full, int bnx, int bny, int bnz, int ioff, int joff, int koff,
           int fsx, int fsy, std::vector<double> &vals) {
  // flat_of_id (immersed_operators.hpp:397-447): only box points are indexed.
  int n = A.nrows;
  std::vector<long long> flat_of_id((size_t)n + 1, -1);
  for (int k = 0; k < bnz; ++k)
    for (int j = 0; j < bny; ++j)
      for (int i = 0; i < bnx; ++i) {
        long long fl = ((long long)(k + koff) * fsy + (j + joff)) * fsx + (i + ioff);
        int id = id_full[(size_t)fl];
        if (id > 0) flat_of_id[(size_t)id] = fl;
      }
  size_t nrows = (size_t)bnx * bny * bnz;
  vals.assign(7 * nrows, 0.0);
  long long plane = (long long)fsx * fsy;
  ORC_PAR2
  for (int k = 0; k < bnz; ++k)
    for (int j = 0; j < bny; ++j)
      for (int i = 0; i < bnx; ++i) {
        size_t row = ((size_t)k * bny + j) * bnx + i;
        long long fl = ((long long)(k + koff) * fsy + (j + joff)) * fsx + (i + ioff);
        int aid = id_full[(size_t)fl];
        double *a = &vals[7 * row];
        if (aid <= 0) { a[0] = 1.0; continue; }
        for (const auto &kv : A.rows[(size_t)(aid - 1)]) {
          long long f2 = flat_of_id[(size_t)kv.first];
          long long kc = f2 / plane, rem = f2 - kc * plane, jc = rem / fsx, ic = rem - jc * fsx;
          long long di = ic - (i + ioff), dj = jc - (j + joff), dk = kc - (k + koff);
          int s = -1;
          if (di == 0 && dj == 0 && dk == 0) s = 0;
          else if (di == -1 && dj == 0 && dk == 0) s = 1;
          else if (di == 1 && dj == 0 && dk == 0) s = 2;
          else if (di == 0 && dj == -1 && dk == 0) s = 3;
          else if (di == 0 && dj == 1 && dk == 0) s = 4;
          else if (di == 0 && dj == 0 && dk == -1) s = 5;
          else if (di == 0 && dj == 0 && dk == 1) s = 6;
          if (s >= 0) a[s] += kv.second;  // off-stencil entries are dropped, as in the reference
        }
      }
}

Setup *setup_create(const orc_params &par) {
  Setup *S = new Setup();
  S->par = par;
  S->nx = par.nx; S->ny = par.ny; S->nz = par.nz;
  S->dx = par.Lx / par.nx; S->dy = par.Ly / par.ny; S->dz = par.Lz / par.nz;
  geom_beta(*S);
  geom_alpha(*S);
  geom_prune(*S);
  extrude_and_ids(*S);
  Sp Au0, Av0, Aw0, Ap0, Dm, Gm, Au, Av, Aw;
  // the four assemblies are independent (each one's rows are filled in its own
  // serial order); the row-parallel helpers inside run serially within a section
#pragma omp parallel sections num_threads(g_threads > 1 ? 4 : 1)
  {
#pragma omp section
    { build_u(*S, Au0, S->rhs_u_bc); sp_row_scale(Au0, Au, S->scale_u); }
#pragma omp section
    { build_v(*S, Av0, S->rhs_v_bc); sp_row_scale(Av0, Av, S->scale_v); }
#pragma omp section
    { build_w(*S, Aw0, S->rhs_w_bc); sp_row_scale(Aw0, Aw, S->scale_w); }
#pragma omp section
    build_proj(*S, Dm, Gm, S->div_bc);
  }
  Ap0 = sp_mul(Dm, Gm);
  for (auto &r : Ap0.rows)
    for (auto &kv : r) kv.second = -kv.second;  // main.cu:101-103
  int nx = S->nx, ny = S->ny, nz = S->nz;
  embed(Au, S->id_u, nx - 1, ny, nz, 1, 0, 0, nx + 1, ny, S->Au_vals);
  embed(Av, S->id_v, nx, ny - 1, nz, 0, 1, 0, nx, ny + 1, S->Av_vals);
  embed(Aw, S->id_w, nx, ny, nz - 1, 0, 0, 1, nx, ny, S->Aw_vals);
  embed(Ap0, S->id_p, nx, ny, nz, 0, 0, 0, nx, ny, S->Ap_vals);
  S->D = sp_to_csr(Dm);
  S->G = sp_to_csr(Gm);
  for (Sp *A : {&Au0, &Av0, &Aw0, &Ap0, &Dm, &Gm, &Au, &Av, &Aw}) sp_release(*A);
  S->uin.assign((size_t)ny * nz, 0.0);  // immersed_operators.hpp:921-925
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      S->uin[(size_t)j + (size_t)ny * k] = inlet_profile((j + 0.5) * S->dy, (k + 0.5) * S->dz, par.Ly, par.Um);
  return S;
}

// ---------------------------------------------------------------- box operator
struct Box {
  int nx = 0, ny = 0, nz = 0;
  std::vector<double> a[7];  // SoA planes, entry order C,-x,+x,-y,+y,-z,+z
  size_t n() const { return (size_t)nx * ny * nz; }
};

Box box_from_aos(int nx, int ny, int nz, const double *vals) {
  Box B;
  B.nx = nx; B.ny = ny; B.nz = nz;
  size_t N = B.n();
  for (int s = 0; s < 7; ++s) {
    B.a[s].resize(N);
    for (size_t r = 0; r < N; ++r) B.a[s][r] = vals[7 * r + s];
  }
  return B;
}

// Neighbour offset of stencil entry s in a box; returns false when out of box.
inline bool nbr(const Box &B, int i, int j, int k, int s, size_t &r) {
  int ii = i, jj = j, kk = k;
  switch (s) {
    case 1: ii--; break; case 2: ii++; break; case 3: jj--; break;
    case 4: jj++; break; case 5: kk--; break; case 6: kk++; break;
    default: break;
  }
  if (ii < 0 || ii >= B.nx || jj < 0 || jj >= B.ny || kk < 0 || kk >= B.nz) return false;
  r = ((size_t)kk * B.ny + jj) * B.nx + ii;
  return true;
}

// y = A x; couplings that leave the box contribute nothing (Struct semantics).
void box_matvec(const Box &B, const double *x, double *y) {
  ORC_PAR2
  for (int k = 0; k < B.nz; ++k)
    for (int j = 0; j < B.ny; ++j)
      for (int i = 0; i < B.nx; ++i) {
        size_t r = ((size_t)k * B.ny + j) * B.nx + i, q;
        double s = B.a[0][r] * x[r];
        for (int e = 1; e < 7; ++e)
          if (nbr(B, i, j, k, e, q)) s += B.a[e][r] * x[q];
        y[r] = s;
      }
}

// ---------------------------------------------------------------- PFMG-style MG
// Definition used by both this oracle and the GPU path (HYPRE's exact PFMG
// internals are unpinned):
//  * level l coarsens by 2 in one direction d_l: the direction with the largest
//    total in-box coupling  c_d = sum_rows |a_-d| + |a_+d|  (rows in k,j,i order;
//    a later direction must exceed the current best by a relative 1e-9);
//    coarse point c sits on fine index 2c+1, nc = floor(n/2);
//  * interpolation at an F point f: w-/+ = -a_-/+d(f) / center(f), where center
//    collapses the non-d couplings onto the diagonal; C points inject; R = P^T;
//  * coarse operator (RAP type 1 analogue, non-Galerkin 7-point): couplings across
//    d are the Galerkin couplings with that cross offset summed over the offsets
//    along d (exact for error that is smooth along d); couplings along d are the
//    1-D Galerkin product of the collapsed stencil (the centre of an F point
//    includes its couplings across d, as in the interpolation); the centre keeps
//    the Galerkin row sum.  For the 7-point Laplacian this is twice the operator
//    rediscretised on the semicoarsened grid.  (Lumping the off-pattern
//    couplings onto the d offsets instead leaves the operator 3/4 too weak for
//    smooth error, and the cycle diverges once the hierarchy is deep.)  RAP
//    type 0 selects the same operator;
//  * relaxation: weighted Jacobi with w_l = 2/(2 + c_d/sum c) (type 1), plain
//    Jacobi (type 0), red/black Gauss-Seidel (types 2: symmetric, 3: forward);
//    one pre and one post sweep;
//  * coarsening stops when a level has at most coarse_max(N0) points, N0 the
//    finest level's, or when no direction has 2+ points; the coarsest system is
//    solved exactly (LU with partial pivoting).  coarse_max(N0) =
//    min(1152, max(64, N0 / 2048)): on the production pressure box (512x96x96)
//    stopping at <= 1152 points takes 13 instead of 18 levels with the same 21
//    GMRES iterations (256x48x48 at <= 1200: 18 and 18), while small boxes need
//    the deep hierarchy (48x16x12 stopped at ~1000 points: 57 instead of 17
//    iterations).  Boxes up to 128x32x32 keep 64: there a larger exact solve
//    would put its own rounding (~cond * eps of the coarsest matrix) at the
//    level of a 1e-12 GMRES tolerance.
inline size_t mg_coarse_max(size_t n0) { return std::min<size_t>(1152, std::max<size_t>(64, n0 / 2048)); }
// A coarsest size set by orc_set_coarse_size (0: the rule above).
static size_t g_coarse = 0;

struct MGLevel {
  Box A;
  int cdir = -1;
  double omega = 1.0;
  std::vector<double> wm, wp;  // interpolation weights at F points (fine level l)
  std::vector<double> x, b, r;
};

struct MG {
  std::vector<MGLevel> L;
  std::vector<double> lu;  // coarsest dense LU (row-major), with pivots
  std::vector<int> piv;
  int relax_type = 1;
};

inline int em(int d) { return 1 + 2 * d; }
inline int ep(int d) { return 2 + 2 * d; }

double coef_in(const Box &B, int i, int j, int k, int s) {
  size_t r = ((size_t)k * B.ny + j) * B.nx + i, q;
  if (s == 0) return B.a[0][r];
  if (!nbr(B, i, j, k, s, q)) return 0.0;
  return B.a[s][r];
}

void mg_choose(const Box &B, int &cdir, double &omega) {
  double c[3] = {0, 0, 0};
  for (int k = 0; k < B.nz; ++k)
    for (int j = 0; j < B.ny; ++j)
      for (int i = 0; i < B.nx; ++i)
        for (int d = 0; d < 3; ++d) c[d] += std::abs(coef_in(B, i, j, k, em(d))) + std::abs(coef_in(B, i, j, k, ep(d)));
  int n[3] = {B.nx, B.ny, B.nz};
  cdir = -1;
  for (int d = 0; d < 3; ++d) {
    if (n[d] < 2) continue;
    if (cdir < 0 || c[d] > c[cdir] * (1.0 + 1.0e-9)) cdir = d;
  }
  double tot = c[0] + c[1] + c[2];
  double s = (cdir >= 0 && tot > 0.0) ? c[cdir] / tot : 1.0;
  omega = 2.0 / (2.0 + s);
}

void mg_weights(MGLevel &Lv) {
  const Box &B = Lv.A;
  int d = Lv.cdir;
  size_t N = B.n();
  Lv.wm.assign(N, 0.0);
  Lv.wp.assign(N, 0.0);
  for (int k = 0; k < B.nz; ++k)
    for (int j = 0; j < B.ny; ++j)
      for (int i = 0; i < B.nx; ++i) {
        int id[3] = {i, j, k};
        if (id[d] % 2 != 0) continue;  // C points (odd) inject
        size_t r = ((size_t)k * B.ny + j) * B.nx + i;
        double ctr = B.a[0][r];
        for (int s = 1; s < 7; ++s)
          if (s != em(d) && s != ep(d)) ctr += coef_in(B, i, j, k, s);
        if (ctr != 0.0) {
          Lv.wm[r] = -coef_in(B, i, j, k, em(d)) / ctr;
          Lv.wp[r] = -coef_in(B, i, j, k, ep(d)) / ctr;
        }
      }
}

Box mg_coarsen(const MGLevel &Lv) {
  const Box &F = Lv.A;
  int d = Lv.cdir;
  Box C;
  int nf[3] = {F.nx, F.ny, F.nz}, nc[3] = {F.nx, F.ny, F.nz};
  nc[d] = nf[d] / 2;
  C.nx = nc[0]; C.ny = nc[1]; C.nz = nc[2];
  size_t Nc = C.n();
  for (int s = 0; s < 7; ++s) C.a[s].assign(Nc, 0.0);
  auto fidx = [&](const int *p) { return ((size_t)p[2] * F.ny + p[1]) * F.nx + p[0]; };
  auto a = [&](const int *p, int s) { return coef_in(F, p[0], p[1], p[2], s); };
  auto inb = [&](const int *p) {
    return p[0] >= 0 && p[0] < F.nx && p[1] >= 0 && p[1] < F.ny && p[2] >= 0 && p[2] < F.nz;
  };
  for (int k = 0; k < C.nz; ++k)
    for (int j = 0; j < C.ny; ++j)
      for (int i = 0; i < C.nx; ++i) {
        int c3[3] = {i, j, k};
        int f0[3] = {i, j, k};
        f0[d] = 2 * c3[d] + 1;
        int fm[3] = {f0[0], f0[1], f0[2]}, fp[3] = {f0[0], f0[1], f0[2]};
        fm[d] -= 1;
        fp[d] += 1;
        bool hasp = inb(fp);
        double rm = Lv.wp[fidx(fm)];
        double rp = hasp ? Lv.wm[fidx(fp)] : 0.0;
        double a0m = a(fm, 0), a0p = hasp ? a(fp, 0) : 0.0;
        size_t rc = ((size_t)k * C.ny + j) * C.nx + i;
        // centre
        double ac = a(f0, 0) + rm * (a(f0, em(d)) + a(fm, ep(d)) + rm * a0m);
        if (hasp) ac += rp * (a(f0, ep(d)) + a(fp, em(d)) + rp * a0p);
        // along d: the 1-D Galerkin product of the collapsed stencil (centre of an
        // F point = its diagonal plus its couplings across d)
        double wmm = Lv.wm[fidx(fm)];
        double amG = a(f0, em(d)) * wmm + rm * a0m * wmm + rm * a(fm, em(d));
        double cmm = a0m, cpp = a0p;
        for (int s = 1; s < 7; ++s) {
          if (s == em(d) || s == ep(d)) continue;
          cmm += a(fm, s);
          if (hasp) cpp += a(fp, s);
        }
        double am_ = a(f0, em(d)) * wmm + rm * cmm * wmm + rm * a(fm, em(d));
        double apG = 0.0, ap_ = 0.0;
        if (hasp) {
          double wpp = Lv.wp[fidx(fp)];
          apG = a(f0, ep(d)) * wpp + rp * a0p * wpp + rp * a(fp, ep(d));
          ap_ = a(f0, ep(d)) * wpp + rp * cpp * wpp + rp * a(fp, ep(d));
        }
        // across d: every Galerkin coupling with this cross offset, summed over the
        // offsets along d
        for (int s = 1; s < 7; ++s) {
          if (s == em(d) || s == ep(d)) continue;
          int e = (s - 1) / 2, sg = (s % 2 == 1) ? -1 : 1;
          double v = a(f0, s);
          double cm = a(fm, s);
          if (cm != 0.0) {
            int q[3] = {fm[0], fm[1], fm[2]};
            q[e] += sg;
            size_t fq = fidx(q);
            v += rm * cm * (Lv.wm[fq] + Lv.wp[fq]);
          }
          if (hasp) {
            double cp = a(fp, s);
            if (cp != 0.0) {
              int q[3] = {fp[0], fp[1], fp[2]};
              q[e] += sg;
              size_t fq = fidx(q);
              v += rp * cp * (Lv.wm[fq] + Lv.wp[fq]);
            }
          }
          C.a[s][rc] = v;
        }
        // centre: keeps the Galerkin row sum
        C.a[0][rc] = ac + (amG - am_) + (apG - ap_);
        C.a[em(d)][rc] = am_;
        C.a[ep(d)][rc] = ap_;
      }
  return C;
}

// Dense LU with partial pivoting, PA = LU, of a matrix whose nonzeros lie within
// bw of the diagonal (7-point box in x-fastest order: bw = nx*ny, or nx, or 1).
// Below the pivot only rows c+1..c+bw can be nonzero in column c, and row swaps
// and updates reach at most column c+2bw, so the loops stop there: the result is
// the same as the full dense elimination (the skipped terms are exact zeros).
void lu_factor(int n, int bw, std::vector<double> &A, std::vector<int> &piv) {
  piv.assign((size_t)n, 0);
  for (int c = 0; c < n; ++c) {
    const int rmax = std::min(n - 1, c + bw), tmax = std::min(n - 1, c + 2 * bw);
    int p = c;
    for (int r = c + 1; r <= rmax; ++r)
      if (std::abs(A[(size_t)r * n + c]) > std::abs(A[(size_t)p * n + c])) p = r;
    piv[(size_t)c] = p;
    if (p != c)
      for (int t = 0; t < n; ++t) std::swap(A[(size_t)c * n + t], A[(size_t)p * n + t]);
    double d = A[(size_t)c * n + c];
    if (d == 0.0) continue;
    for (int r = c + 1; r <= rmax; ++r) {
      double f = A[(size_t)r * n + c] / d;
      A[(size_t)r * n + c] = f;
      for (int t = c + 1; t <= tmax; ++t) A[(size_t)r * n + t] -= f * A[(size_t)c * n + t];
    }
  }
}

void lu_solve(int n, const std::vector<double> &A, const std::vector<int> &piv, double *x) {
  for (int c = 0; c < n; ++c) {
    int p = piv[(size_t)c];
    if (p != c) std::swap(x[c], x[p]);
    for (int r = c + 1; r < n; ++r) x[r] -= A[(size_t)r * n + c] * x[c];
  }
  for (int r = n - 1; r >= 0; --r) {
    double s = x[r];
    for (int t = r + 1; t < n; ++t) s -= A[(size_t)r * n + t] * x[t];
    double d = A[(size_t)r * n + r];
    x[r] = (d != 0.0) ? s / d : 0.0;
  }
}

void mg_setup(MG &M, const Box &A0, int relax_type, int max_levels) {
  M.relax_type = relax_type;
  M.L.clear();
  M.L.emplace_back();
  M.L.back().A = A0;
  const size_t cmax = g_coarse > 0 ? g_coarse : mg_coarse_max(A0.n());
  for (;;) {
    MGLevel &Lv = M.L.back();
    mg_choose(Lv.A, Lv.cdir, Lv.omega);
    size_t N = Lv.A.n();
    Lv.x.assign(N, 0.0); Lv.b.assign(N, 0.0); Lv.r.assign(N, 0.0);
    if (N <= cmax || Lv.cdir < 0 || (max_levels > 0 && (int)M.L.size() >= max_levels)) {
      Lv.cdir = -1;
      break;
    }
    mg_weights(Lv);
    Box C = mg_coarsen(Lv);
    M.L.emplace_back();
    M.L.back().A = std::move(C);
  }
  const Box &B = M.L.back().A;
  int n = (int)B.n();
  M.lu.assign((size_t)n * n, 0.0);
  for (int k = 0; k < B.nz; ++k)
    for (int j = 0; j < B.ny; ++j)
      for (int i = 0; i < B.nx; ++i) {
        size_t r = ((size_t)k * B.ny + j) * B.nx + i, q;
        M.lu[r * n + r] += B.a[0][r];
        for (int s = 1; s < 7; ++s)
          if (nbr(B, i, j, k, s, q)) M.lu[r * n + q] += B.a[s][r];
      }
  const int bw = (B.nz > 1) ? B.nx * B.ny : (B.ny > 1 ? B.nx : 1);
  lu_factor(n, bw, M.lu, M.piv);
}

// One relaxation sweep on level l.  zero: x is taken as 0 on entry.
void mg_relax(MG &M, MGLevel &Lv, bool zero, bool post) {
  const Box &B = Lv.A;
  size_t N = B.n();
  if (M.relax_type <= 1) {
    double w = (M.relax_type == 0) ? 1.0 : Lv.omega;
    if (zero) {
      ORC_PAR
      for (size_t r = 0; r < N; ++r) Lv.x[r] = (B.a[0][r] != 0.0) ? w * Lv.b[r] / B.a[0][r] : 0.0;
      return;
    }
    box_matvec(B, Lv.x.data(), Lv.r.data());
    ORC_PAR
    for (size_t r = 0; r < N; ++r)
      if (B.a[0][r] != 0.0) Lv.x[r] = Lv.x[r] + w * (Lv.b[r] - Lv.r[r]) / B.a[0][r];
    return;
  }
  if (zero) std::fill(Lv.x.begin(), Lv.x.end(), 0.0);
  int order[2] = {0, 1};
  if (post && M.relax_type == 2) { order[0] = 1; order[1] = 0; }
  for (int pass = 0; pass < 2; ++pass) {
    int col = order[pass];
    ORC_PAR2
    for (int k = 0; k < B.nz; ++k)
      for (int j = 0; j < B.ny; ++j)
        for (int i = 0; i < B.nx; ++i) {
          if (((i + j + k) & 1) != col) continue;
          size_t r = ((size_t)k * B.ny + j) * B.nx + i, q;
          if (B.a[0][r] == 0.0) continue;
          double s = Lv.b[r];
          for (int e = 1; e < 7; ++e)
            if (nbr(B, i, j, k, e, q)) s -= B.a[e][r] * Lv.x[q];
          Lv.x[r] = s / B.a[0][r];
        }
  }
}

// One V-cycle with zero initial guess: x = M^-1 b.
void mg_vcycle(MG &M, const double *b, double *x) {
  int nl = (int)M.L.size();
  std::copy(b, b + M.L[0].A.n(), M.L[0].b.begin());
  for (int l = 0; l < nl - 1; ++l) {
    MGLevel &F = M.L[(size_t)l];
    MGLevel &C = M.L[(size_t)l + 1];
    mg_relax(M, F, true, false);
    box_matvec(F.A, F.x.data(), F.r.data());
    for (size_t r = 0; r < F.A.n(); ++r) F.r[r] = F.b[r] - F.r[r];
    int d = F.cdir;
    ORC_PAR2
    for (int k = 0; k < C.A.nz; ++k)
      for (int j = 0; j < C.A.ny; ++j)
        for (int i = 0; i < C.A.nx; ++i) {
          int f0[3] = {i, j, k};
          f0[d] = 2 * f0[d] + 1;
          int fm[3] = {f0[0], f0[1], f0[2]}, fp[3] = {f0[0], f0[1], f0[2]};
          fm[d]--;
          fp[d]++;
          int nf[3] = {F.A.nx, F.A.ny, F.A.nz};
          auto fi = [&](const int *p) { return ((size_t)p[2] * F.A.ny + p[1]) * F.A.nx + p[0]; };
          double v = F.r[fi(f0)] + F.wp[fi(fm)] * F.r[fi(fm)];
          if (fp[d] < nf[d]) v += F.wm[fi(fp)] * F.r[fi(fp)];
          C.b[((size_t)k * C.A.ny + j) * C.A.nx + i] = v;
        }
  }
  MGLevel &Lc = M.L.back();
  std::copy(Lc.b.begin(), Lc.b.end(), Lc.x.begin());
  lu_solve((int)Lc.A.n(), M.lu, M.piv, Lc.x.data());
  for (int l = nl - 2; l >= 0; --l) {
    MGLevel &F = M.L[(size_t)l];
    MGLevel &C = M.L[(size_t)l + 1];
    int d = F.cdir;
    int nc[3] = {C.A.nx, C.A.ny, C.A.nz};
    ORC_PAR2
    for (int k = 0; k < F.A.nz; ++k)
      for (int j = 0; j < F.A.ny; ++j)
        for (int i = 0; i < F.A.nx; ++i) {
          int p[3] = {i, j, k};
          size_t r = ((size_t)k * F.A.ny + j) * F.A.nx + i;
          auto ci = [&](const int *q) { return ((size_t)q[2] * C.A.ny + q[1]) * C.A.nx + q[0]; };
          if (p[d] % 2 == 1) {
            int q[3] = {p[0], p[1], p[2]};
            q[d] = (p[d] - 1) / 2;
            F.x[r] += C.x[ci(q)];
          } else {
            int cL = p[d] / 2 - 1, cR = p[d] / 2;
            double t = 0.0;
            if (cL >= 0) { int q[3] = {p[0], p[1], p[2]}; q[d] = cL; t += F.wm[r] * C.x[ci(q)]; }
            if (cR < nc[d]) { int q[3] = {p[0], p[1], p[2]}; q[d] = cR; t += F.wp[r] * C.x[ci(q)]; }
            F.x[r] += t;
          }
        }
    mg_relax(M, F, false, true);
  }
  std::copy(M.L[0].x.begin(), M.L[0].x.end(), x);
}

// ---------------------------------------------------------------- GMRES (HYPRE control flow)
struct Precond {
  int kind = 0;  // 0 none, 1 Jacobi, 2 MG
  int iters = 1;
  const Box *A = nullptr;
  MG mg;
  void apply(const double *b, double *x) {  // x = M^-1 b, zero initial guess
    size_t N = A->n();
    if (kind == 0) { std::copy(b, b + N, x); return; }
    if (kind == 1) {
      ORC_PAR
      for (size_t r = 0; r < N; ++r) x[r] = b[r] / A->a[0][r];
      std::vector<double> t(N);
      for (int it = 1; it < iters; ++it) {
        box_matvec(*A, x, t.data());
        ORC_PAR
        for (size_t r = 0; r < N; ++r) x[r] = x[r] + (b[r] - t[r]) / A->a[0][r];
      }
      return;
    }
    mg_vcycle(mg, b, x);
    if (iters > 1) {
      std::vector<double> r(N), e(N);
      for (int it = 1; it < iters; ++it) {
        box_matvec(*A, x, r.data());
        ORC_PAR
        for (size_t q = 0; q < N; ++q) r[q] = b[q] - r[q];
        mg_vcycle(mg, r.data(), e.data());
        ORC_PAR
        for (size_t q = 0; q < N; ++q) x[q] += e[q];
      }
    }
  }
};

double dot(const std::vector<double> &a, const std::vector<double> &b) {
  const size_t n = a.size();
  if (g_threads <= 1) {
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
  }
  std::vector<double> part((size_t)g_threads, 0.0);
  int team = 1;
#pragma omp parallel num_threads(g_threads)
  {
    const int nt = omp_get_num_threads(), t = omp_get_thread_num();
    const size_t lo = n * (size_t)t / (size_t)nt, hi = n * ((size_t)t + 1) / (size_t)nt;
    double s = 0.0;
    for (size_t i = lo; i < hi; ++i) s += a[i] * b[i];
    part[(size_t)t] = s;
    if (t == 0) team = nt;
  }
  double s = 0.0;
  for (int t = 0; t < team; ++t) s += part[(size_t)t];
  return s;
}

int gmres_solve(const Box &A, Precond &M, const double *bp, double *xp, double tol, int maxit, int kdim,
                int *its_out, double *rel_out) {
  size_t N = A.n();
  std::vector<double> b(bp, bp + N), x(xp, xp + N), r(N), w(N);
  std::vector<std::vector<double>> p((size_t)kdim + 1, std::vector<double>(N));
  std::vector<std::vector<double>> hh((size_t)kdim + 1, std::vector<double>((size_t)kdim, 0.0));
  std::vector<double> c((size_t)kdim), s((size_t)kdim), rs((size_t)kdim + 1);
  auto matvec_res = [&](std::vector<double> &out) {  // out = b - A x
    box_matvec(A, x.data(), out.data());
    ORC_PAR
    for (size_t q = 0; q < N; ++q) out[q] = b[q] - out[q];
  };
  matvec_res(p[0]);
  double b_norm = std::sqrt(dot(b, b));
  double r_norm = std::sqrt(dot(p[0], p[0]));
  if (!std::isfinite(b_norm) || !std::isfinite(r_norm)) { *its_out = 0; *rel_out = NAN; return 1; }
  double den = (b_norm > 0.0) ? b_norm : r_norm;
  double eps = tol * den;
  double real_old = b_norm;
  int iter = 0, i = 0;
  bool conv = false;
  while (iter < maxit) {
    rs[0] = r_norm;
    if (r_norm == 0.0) { conv = true; break; }
    if (r_norm <= eps) {
      matvec_res(r);
      r_norm = std::sqrt(dot(r, r));
      if (r_norm <= eps) { conv = true; break; }
    }
    double t = 1.0 / r_norm;
    ORC_PAR
    for (size_t q = 0; q < N; ++q) p[0][q] *= t;
    i = 0;
    while (i < kdim && iter < maxit) {
      i++;
      iter++;
      M.apply(p[(size_t)i - 1].data(), r.data());
      box_matvec(A, r.data(), p[(size_t)i].data());
      for (int j = 0; j < i; ++j) {  // modified Gram-Schmidt
        hh[(size_t)j][(size_t)i - 1] = dot(p[(size_t)j], p[(size_t)i]);
        double h = hh[(size_t)j][(size_t)i - 1];
        ORC_PAR
        for (size_t q = 0; q < N; ++q) p[(size_t)i][q] -= h * p[(size_t)j][q];
      }
      t = std::sqrt(dot(p[(size_t)i], p[(size_t)i]));
      hh[(size_t)i][(size_t)i - 1] = t;
      if (t != 0.0) {
        t = 1.0 / t;
        ORC_PAR
        for (size_t q = 0; q < N; ++q) p[(size_t)i][q] *= t;
      }
      for (int j = 1; j < i; ++j) {
        double tt = hh[(size_t)j - 1][(size_t)i - 1];
        hh[(size_t)j - 1][(size_t)i - 1] = s[(size_t)j - 1] * hh[(size_t)j][(size_t)i - 1] + c[(size_t)j - 1] * tt;
        hh[(size_t)j][(size_t)i - 1] = -s[(size_t)j - 1] * tt + c[(size_t)j - 1] * hh[(size_t)j][(size_t)i - 1];
      }
      double g = std::sqrt(hh[(size_t)i][(size_t)i - 1] * hh[(size_t)i][(size_t)i - 1] +
                           hh[(size_t)i - 1][(size_t)i - 1] * hh[(size_t)i - 1][(size_t)i - 1]);
      if (g == 0.0) g = std::numeric_limits<double>::epsilon();
      c[(size_t)i - 1] = hh[(size_t)i - 1][(size_t)i - 1] / g;
      s[(size_t)i - 1] = hh[(size_t)i][(size_t)i - 1] / g;
      rs[(size_t)i] = -hh[(size_t)i][(size_t)i - 1] * rs[(size_t)i - 1] / g;
      rs[(size_t)i - 1] = c[(size_t)i - 1] * rs[(size_t)i - 1];
      hh[(size_t)i - 1][(size_t)i - 1] = s[(size_t)i - 1] * hh[(size_t)i][(size_t)i - 1] + c[(size_t)i - 1] * hh[(size_t)i - 1][(size_t)i - 1];
      r_norm = std::abs(rs[(size_t)i]);
      if (r_norm <= eps) break;
    }
    // back substitution
    rs[(size_t)i - 1] = rs[(size_t)i - 1] / hh[(size_t)i - 1][(size_t)i - 1];
    for (int k = i - 2; k >= 0; --k) {
      double tt = 0.0;
      for (int j = k + 1; j < i; ++j) tt -= hh[(size_t)k][(size_t)j] * rs[(size_t)j];
      tt += rs[(size_t)k];
      rs[(size_t)k] = tt / hh[(size_t)k][(size_t)k];
    }
    ORC_PAR
    for (size_t q = 0; q < N; ++q) w[q] = rs[(size_t)i - 1] * p[(size_t)i - 1][q];
    for (int j = i - 2; j >= 0; --j)
      ORC_PAR
      for (size_t q = 0; q < N; ++q) w[q] += rs[(size_t)j] * p[(size_t)j][q];
    M.apply(w.data(), r.data());
    ORC_PAR
    for (size_t q = 0; q < N; ++q) x[q] += r[q];
    if (r_norm <= eps) {
      matvec_res(r);
      double real_new = std::sqrt(dot(r, r));
      r_norm = real_new;
      if (r_norm <= eps) { conv = true; break; }
      if (real_new >= real_old) { conv = true; break; }
      p[0] = r;
      i = 0;
      real_old = real_new;
    }
    // residual vector for the restart from the Givens-rotated basis
    for (int j = i; j > 0; --j) {
      rs[(size_t)j - 1] = -s[(size_t)j - 1] * rs[(size_t)j];
      rs[(size_t)j] = c[(size_t)j - 1] * rs[(size_t)j];
    }
    if (i) {
      ORC_PAR
      for (size_t q = 0; q < N; ++q) p[(size_t)i][q] += (rs[(size_t)i] - 1.0) * p[(size_t)i][q];
      for (int j = i - 1; j > 0; --j)
        ORC_PAR
        for (size_t q = 0; q < N; ++q) p[(size_t)i][q] += rs[(size_t)j] * p[(size_t)j][q];
      ORC_PAR
      for (size_t q = 0; q < N; ++q) p[0][q] += (rs[0] - 1.0) * p[0][q];
      ORC_PAR
      for (size_t q = 0; q < N; ++q) p[0][q] += p[(size_t)i][q];
    }
  }
  std::copy(x.begin(), x.end(), xp);
  *its_out = iter;
  *rel_out = (b_norm > 0.0) ? r_norm / b_norm : r_norm;
  if (!conv && iter >= maxit && r_norm > eps && eps > 0) return 256;
  return 0;
}

// ---------------------------------------------------------------- simulation
uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// Seeded uniform noise in (-amp, amp) for field f (0 u, 1 v, 2 w) at flat index idx.
double noise_at(uint64_t seed, int f, size_t idx, double amp) {
  uint64_t key = splitmix64(seed ^ ((uint64_t)(f + 1) << 56));
  uint64_t z = splitmix64(key + (uint64_t)idx);
  double u01 = (double)(z >> 11) * (1.0 / 9007199254740992.0);
  return amp * (2.0 * u01 - 1.0);
}

struct Sim {
  Setup *S = nullptr;
  Box Bu, Bv, Bw, Bp;
  Precond Mu, Mv, Mw, Mp;
  int step = 0;
  double ramp = 0.0;
  std::vector<double> u, v, w, p, ut, vt, wt, uin_step;
  std::vector<double> uin_base, div_bc_base;  // before an inlet amplitude field was applied
  std::vector<double> conv_u, conv_v, conv_w, gpx, gpy, gpz;
  std::vector<double> rhs_u, x_u, rhs_v, x_v, rhs_w, x_w, rhs_p, x_p;
  std::vector<double> pvec, gp, qt, Dqt, phi, gphi, divv;
  int its[4] = {0, 0, 0, 0};
  double rel[4] = {0, 0, 0, 0};
  int err[4] = {0, 0, 0, 0};
  double div_l2 = 0, umax = 0, vmax = 0, wmax = 0, fx = 0, fy = 0, cd = 0, cl = 0;
  // force functional weights (see forces())
  std::vector<std::pair<size_t, double>> fu_w, fv_w;
  std::vector<std::pair<size_t, double>> px_w, py_w;
};

// BC cleanup with the GPU kernel's element semantics (gpu_dense_ops.hpp:5-10).
// The outflow copy reads the already-cleaned last interior face.
void bc_cleanup(Sim &M, std::vector<double> &u, std::vector<double> &v, std::vector<double> &w) {
  const Setup &S = *M.S;
  int nx = S.nx, ny = S.ny, nz = S.nz;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        size_t t = ixU(i, j, k, nx, ny);
        if (i == 0) u[t] = M.uin_step[(size_t)j + (size_t)ny * k];
        else if (i < nx && !S.active_u[t]) u[t] = 0.0;
      }
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) u[ixU(nx, j, k, nx, ny)] = u[ixU(nx - 1, j, k, nx, ny)];
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t t = ixV(i, j, k, nx, ny);
        if (j == 0 || j == ny || !S.active_v[t]) v[t] = 0.0;
      }
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t t = ixW(i, j, k, nx, ny);
        if (k == 0 || k == nz) w[t] = 0.0;
        else if (!S.active_w[ixW(i, j, k - 1, nx, ny)]) w[t] = 0.0;  // reference tests active_w[k-1]
      }
}

// Centred convection (gpu_dense_ops.hpp:1-3, 12-14).
void convection(Sim &M) {
  const Setup &S = *M.S;
  int nx = S.nx, ny = S.ny, nz = S.nz;
  double dx = S.dx, dy = S.dy, dz = S.dz;
  const std::vector<double> &u = M.u, &v = M.v, &w = M.w;
  auto U = [&](int i, int j, int k, double fb) { return (i < 0 || i > nx || j < 0 || j >= ny || k < 0 || k >= nz) ? fb : u[ixU(i, j, k, nx, ny)]; };
  auto V = [&](int i, int j, int k, double fb) { return (i < 0 || i >= nx || j < 0 || j > ny || k < 0 || k >= nz) ? fb : v[ixV(i, j, k, nx, ny)]; };
  auto W = [&](int i, int j, int k, double fb) { return (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k > nz) ? fb : w[ixW(i, j, k, nx, ny)]; };
  M.conv_u.assign(u.size(), 0.0);
  M.conv_v.assign(v.size(), 0.0);
  M.conv_w.assign(w.size(), 0.0);
  ORC_PAR2
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 1; i <= nx - 1; ++i) {
        size_t t = ixU(i, j, k, nx, ny);
        if (!S.active_u[t]) continue;
        double uc = u[t];
        double va = 0.25 * (V(i - 1, j, k, 0.0) + V(i - 1, j + 1, k, 0.0) + V(i, j, k, 0.0) + V(i, j + 1, k, 0.0));
        double wa = 0.25 * (W(i - 1, j, k, 0.0) + W(i - 1, j, k + 1, 0.0) + W(i, j, k, 0.0) + W(i, j, k + 1, 0.0));
        double ue = U(i + 1, j, k, uc), uw = U(i - 1, j, k, uc), un = U(i, j + 1, k, uc), us = U(i, j - 1, k, uc);
        double ut = U(i, j, k + 1, uc), ub = U(i, j, k - 1, uc);
        M.conv_u[t] = uc * (ue - uw) / (2 * dx) + va * (un - us) / (2 * dy) + wa * (ut - ub) / (2 * dz);
      }
  ORC_PAR2
  for (int k = 0; k < nz; ++k)
    for (int j = 1; j <= ny - 1; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t t = ixV(i, j, k, nx, ny);
        if (!S.active_v[t]) continue;
        double vc = v[t];
        double ua = 0.25 * (U(i, j - 1, k, 0.0) + U(i + 1, j - 1, k, 0.0) + U(i, j, k, 0.0) + U(i + 1, j, k, 0.0));
        double wa = 0.25 * (W(i, j - 1, k, 0.0) + W(i, j, k, 0.0) + W(i, j - 1, k + 1, 0.0) + W(i, j, k + 1, 0.0));
        double ve = V(i + 1, j, k, vc), vw = V(i - 1, j, k, vc), vn = V(i, j + 1, k, vc), vs = V(i, j - 1, k, vc);
        double vt = V(i, j, k + 1, vc), vb = V(i, j, k - 1, vc);
        M.conv_v[t] = ua * (ve - vw) / (2 * dx) + vc * (vn - vs) / (2 * dy) + wa * (vt - vb) / (2 * dz);
      }
  ORC_PAR2
  for (int k = 1; k <= nz - 1; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t t = ixW(i, j, k, nx, ny);
        if (!S.active_w[ixW(i, j, k - 1, nx, ny)]) continue;
        double wc = w[t];
        double ua = 0.25 * (U(i, j, k - 1, 0.0) + U(i + 1, j, k - 1, 0.0) + U(i, j, k, 0.0) + U(i + 1, j, k, 0.0));
        double va = 0.25 * (V(i, j, k - 1, 0.0) + V(i, j + 1, k - 1, 0.0) + V(i, j, k, 0.0) + V(i, j + 1, k, 0.0));
        double we = W(i + 1, j, k, wc), ww = W(i - 1, j, k, wc), wn = W(i, j + 1, k, wc), ws = W(i, j - 1, k, wc);
        double wt = W(i, j, k + 1, wc), wb = W(i, j, k - 1, wc);
        M.conv_w[t] = ua * (we - ww) / (2 * dx) + va * (wn - ws) / (2 * dy) + wc * (wt - wb) / (2 * dz);
      }
}

// G p through pack -> CSR -> unpack (main.cu:262-271).
void grad_faces(Sim &M, const std::vector<double> &pfield, std::vector<double> &gx, std::vector<double> &gy,
                std::vector<double> &gz) {
  const Setup &S = *M.S;
  int nx = S.nx, ny = S.ny, nz = S.nz;
  size_t Nq = (size_t)S.Nqu + S.Nqv + S.Nqw;
  M.pvec.assign((size_t)S.Np + 1, 0.0);
  for (size_t c = 0; c < S.fluid_p.size(); ++c)
    if (S.fluid_p[c] && S.id_p[c] > 0) M.pvec[(size_t)S.id_p[c]] = pfield[c];
  M.gp.assign(Nq + 1, 0.0);
  csr_spmv(S.G, M.pvec.data(), M.gp.data());
  gx.assign((size_t)(nx + 1) * ny * nz, 0.0);
  gy.assign((size_t)nx * (ny + 1) * nz, 0.0);
  gz.assign((size_t)nx * ny * (nz + 1), 0.0);
  for (size_t t = 0; t < gx.size(); ++t) { int id = S.id_qu[t]; gx[t] = id > 0 ? M.gp[(size_t)id] : 0.0; }
  for (size_t t = 0; t < gy.size(); ++t) { int id = S.id_qv[t]; gy[t] = id > 0 ? M.gp[(size_t)S.Nqu + id] : 0.0; }
  for (size_t t = 0; t < gz.size(); ++t) { int id = S.id_qw[t]; gz[t] = id > 0 ? M.gp[(size_t)S.Nqu + S.Nqv + id] : 0.0; }
}

void phase_ramp(Sim &M) {
  M.step += 1;
  M.ramp = std::min(1.0, (double)M.step / (double)M.S->par.inlet_ramp_steps);
  for (size_t t = 0; t < M.S->uin.size(); ++t) M.uin_step[t] = M.ramp * M.S->uin[t];
}

// main.cu:253-275
void phase_explicit(Sim &M) {
  const Setup &S = *M.S;
  int nx = S.nx, ny = S.ny, nz = S.nz;
  double rho = S.par.rho, dt = S.par.dt;
  int zero = S.par.use_zero_initial_guess != 0;
  bc_cleanup(M, M.u, M.v, M.w);
  convection(M);
  grad_faces(M, M.p, M.gpx, M.gpy, M.gpz);
  M.rhs_u.assign((size_t)(nx - 1) * ny * nz, 0.0); M.x_u.assign(M.rhs_u.size(), 0.0);
  M.rhs_v.assign((size_t)nx * (ny - 1) * nz, 0.0); M.x_v.assign(M.rhs_v.size(), 0.0);
  M.rhs_w.assign((size_t)nx * ny * (nz - 1), 0.0); M.x_w.assign(M.rhs_w.size(), 0.0);
  ORC_PAR2
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 1; i <= nx - 1; ++i) {
        size_t row = ((size_t)k * ny + j) * (nx - 1) + (i - 1), t = ixU(i, j, k, nx, ny);
        int id = S.id_u[t];
        double r = 0.0;
        if (id > 0 && S.active_u[t]) r = ((rho / dt) * M.u[t] - rho * M.conv_u[t] - M.gpx[t] + S.rhs_u_bc[(size_t)id]) * S.scale_u[(size_t)id];
        M.rhs_u[row] = r;
        M.x_u[row] = zero ? 0.0 : M.u[t];
      }
  ORC_PAR2
  for (int k = 0; k < nz; ++k)
    for (int j = 1; j <= ny - 1; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t row = ((size_t)k * (ny - 1) + (j - 1)) * nx + i, t = ixV(i, j, k, nx, ny);
        int id = S.id_v[t];
        double r = 0.0;
        if (id > 0 && S.active_v[t]) r = ((rho / dt) * M.v[t] - rho * M.conv_v[t] - M.gpy[t] + S.rhs_v_bc[(size_t)id]) * S.scale_v[(size_t)id];
        M.rhs_v[row] = r;
        M.x_v[row] = zero ? 0.0 : M.v[t];
      }
  ORC_PAR2
  for (int k = 1; k <= nz - 1; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t row = ((size_t)(k - 1) * ny + j) * nx + i, t = ixW(i, j, k, nx, ny);
        int id = S.id_w[t];
        double r = 0.0;
        if (id > 0 && S.active_w[t]) r = ((rho / dt) * M.w[t] - rho * M.conv_w[t] - M.gpz[t] + S.rhs_w_bc[(size_t)id]) * S.scale_w[(size_t)id];
        M.rhs_w[row] = r;
        M.x_w[row] = zero ? 0.0 : M.w[t];
      }
}

void phase_momentum(Sim &M) {
  const orc_params &p = M.S->par;
  M.err[0] = gmres_solve(M.Bu, M.Mu, M.rhs_u.data(), M.x_u.data(), p.vel_tol, p.vel_maxit, p.vel_kdim, &M.its[0], &M.rel[0]);
  M.err[1] = gmres_solve(M.Bv, M.Mv, M.rhs_v.data(), M.x_v.data(), p.vel_tol, p.vel_maxit, p.vel_kdim, &M.its[1], &M.rel[1]);
  M.err[2] = gmres_solve(M.Bw, M.Mw, M.rhs_w.data(), M.x_w.data(), p.vel_tol, p.vel_maxit, p.vel_kdim, &M.its[2], &M.rel[2]);
}

// main.cu:307-312 (k_scatter_*, then BC cleanup of the tilde fields)
void phase_scatter(Sim &M) {
  const Setup &S = *M.S;
  int nx = S.nx, ny = S.ny, nz = S.nz;
  M.ut = M.u; M.vt = M.v; M.wt = M.w;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 1; i <= nx - 1; ++i) M.ut[ixU(i, j, k, nx, ny)] = M.x_u[((size_t)k * ny + j) * (nx - 1) + (i - 1)];
  for (int k = 0; k < nz; ++k)
    for (int j = 1; j <= ny - 1; ++j)
      for (int i = 0; i < nx; ++i) M.vt[ixV(i, j, k, nx, ny)] = M.x_v[((size_t)k * (ny - 1) + (j - 1)) * nx + i];
  for (int k = 1; k <= nz - 1; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) M.wt[ixW(i, j, k, nx, ny)] = M.x_w[((size_t)(k - 1) * ny + j) * nx + i];
  bc_cleanup(M, M.ut, M.vt, M.wt);
}

// main.cu:314-321
void phase_divergence(Sim &M) {
  const Setup &S = *M.S;
  size_t Nq = (size_t)S.Nqu + S.Nqv + S.Nqw;
  M.qt.assign(Nq + 1, 0.0);
  for (size_t t = 0; t < M.ut.size(); ++t) { int id = S.id_qu[t]; if (id > 0) M.qt[(size_t)id] = M.ut[t]; }
  for (size_t t = 0; t < M.vt.size(); ++t) { int id = S.id_qv[t]; if (id > 0) M.qt[(size_t)S.Nqu + id] = M.vt[t]; }
  for (size_t t = 0; t < M.wt.size(); ++t) { int id = S.id_qw[t]; if (id > 0) M.qt[(size_t)S.Nqu + S.Nqv + id] = M.wt[t]; }
  M.Dqt.assign((size_t)S.Np + 1, 0.0);
  csr_spmv(S.D, M.qt.data(), M.Dqt.data());
  double rho = S.par.rho, dt = S.par.dt;
  M.rhs_p.assign(S.fluid_p.size(), 0.0);
  M.x_p.assign(S.fluid_p.size(), 0.0);
  for (size_t c = 0; c < S.fluid_p.size(); ++c)
    if (S.fluid_p[c]) {
      int id = S.id_p[c];
      if (id > 0) M.rhs_p[c] = -(rho / dt) * (M.Dqt[(size_t)id] + M.ramp * S.div_bc[(size_t)id]);
    }
}

void phase_pressure(Sim &M) {
  const orc_params &p = M.S->par;
  M.err[3] = gmres_solve(M.Bp, M.Mp, M.rhs_p.data(), M.x_p.data(), p.p_tol, p.p_maxit, p.p_kdim, &M.its[3], &M.rel[3]);
}

// main.cu:338-350
void phase_projection(Sim &M) {
  const Setup &S = *M.S;
  size_t Nq = (size_t)S.Nqu + S.Nqv + S.Nqw;
  M.phi.assign((size_t)S.Np + 1, 0.0);
  for (size_t c = 0; c < S.fluid_p.size(); ++c)
    if (S.fluid_p[c] && S.id_p[c] > 0) M.phi[(size_t)S.id_p[c]] = M.x_p[c];
  M.gphi.assign(Nq + 1, 0.0);
  csr_spmv(S.G, M.phi.data(), M.gphi.data());
  double f = S.par.dt / S.par.rho;
  for (size_t t = 1; t <= Nq; ++t) M.qt[t] -= f * M.gphi[t];
  M.u = M.ut; M.v = M.vt; M.w = M.wt;
  for (size_t t = 0; t < M.u.size(); ++t) { int id = S.id_qu[t]; if (id > 0) M.u[t] = M.qt[(size_t)id]; }
  for (size_t t = 0; t < M.v.size(); ++t) { int id = S.id_qv[t]; if (id > 0) M.v[t] = M.qt[(size_t)S.Nqu + id]; }
  for (size_t t = 0; t < M.w.size(); ++t) { int id = S.id_qw[t]; if (id > 0) M.w[t] = M.qt[(size_t)S.Nqu + S.Nqv + id]; }
  bc_cleanup(M, M.u, M.v, M.w);
  for (size_t c = 0; c < S.fluid_p.size(); ++c)
    if (S.fluid_p[c] && S.id_p[c] > 0) M.p[c] += M.phi[(size_t)S.id_p[c]];
}

// Drag/lift functional (not in the reference):
//   pressure: F_x += p_c (alpha_west - alpha_east) dy dz, F_y += p_c (alpha_south - alpha_north) dx dz
//             over fluid cells (cut cells are the only ones with unequal fractions);
//   viscous : every u (v) momentum closure against the cylinder contributes
//             rho nu A_face u_P / d_P to F_x (F_y), with A_face the cell-face area normal
//             to the blocked direction and d_P the ray distance to the circle;
//   c_D = 2 F_x / (rho Ubar^2 D H), c_L likewise, Ubar = 4 Um / 9, D = 2R, H = Lz.
void force_weights(Sim &M) {
  const Setup &S = *M.S;
  const orc_params &p = S.par;
  int nx = S.nx, ny = S.ny, nz = S.nz;
  double dx = S.dx, dy = S.dy, dz = S.dz, rn = p.rho * p.nu;
  auto seg = [&](double a, double b, double c, double d) { return seg_circle(a, b, c, d, p.xcirc, p.ycirc, p.Rcirc); };
  auto add = [](std::vector<std::pair<size_t, double>> &L, size_t t, double area, double d, double rn_) {
    if (std::isfinite(d) && d > 0.0) L.emplace_back(t, rn_ * area / d);
  };
  M.fu_w.clear(); M.fv_w.clear(); M.px_w.clear(); M.py_w.clear();
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 1; i <= nx - 1; ++i) {
        size_t t = ixU(i, j, k, nx, ny);
        if (!S.active_u[t]) continue;
        double xP = i * dx, yP = (j + 0.5) * dy;
        if (i > 1 && !S.active_u[ixU(i - 1, j, k, nx, ny)]) add(M.fu_w, t, dy * dz, seg(xP, yP, xP - dx, yP), rn);
        if (i < nx - 1 && !S.active_u[ixU(i + 1, j, k, nx, ny)]) add(M.fu_w, t, dy * dz, seg(xP, yP, xP + dx, yP), rn);
        if (j > 0 && !S.active_u[ixU(i, j - 1, k, nx, ny)]) add(M.fu_w, t, dx * dz, seg(xP, yP, xP, yP - dy), rn);
        if (j < ny - 1 && !S.active_u[ixU(i, j + 1, k, nx, ny)]) add(M.fu_w, t, dx * dz, seg(xP, yP, xP, yP + dy), rn);
      }
  for (int k = 0; k < nz; ++k)
    for (int j = 1; j <= ny - 1; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t t = ixV(i, j, k, nx, ny);
        if (!S.active_v[t]) continue;
        double xP = (i + 0.5) * dx, yP = j * dy;
        if (i > 0 && !S.active_v[ixV(i - 1, j, k, nx, ny)]) add(M.fv_w, t, dy * dz, seg(xP, yP, xP - dx, yP), rn);
        if (i < nx - 1 && !S.active_v[ixV(i + 1, j, k, nx, ny)]) add(M.fv_w, t, dy * dz, seg(xP, yP, xP + dx, yP), rn);
        if (j > 1 && !S.active_v[ixV(i, j - 1, k, nx, ny)]) add(M.fv_w, t, dx * dz, seg(xP, yP, xP, yP - dy), rn);
        if (j < ny - 1 && !S.active_v[ixV(i, j + 1, k, nx, ny)]) add(M.fv_w, t, dx * dz, seg(xP, yP, xP, yP + dy), rn);
      }
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t c = ixP(i, j, k, nx, ny);
        if (!S.fluid_p[c]) continue;
        double aw = S.alpha_u2d[(size_t)i + (size_t)j * (nx + 1)], ae = S.alpha_u2d[(size_t)(i + 1) + (size_t)j * (nx + 1)];
        double as = S.alpha_v2d[(size_t)i + (size_t)j * nx], an = S.alpha_v2d[(size_t)i + (size_t)(j + 1) * nx];
        if (aw != ae) M.px_w.emplace_back(c, (aw - ae) * dy * dz);
        if (as != an) M.py_w.emplace_back(c, (as - an) * dx * dz);
      }
}

void phase_diagnostics(Sim &M) {
  const Setup &S = *M.S;
  M.divv.assign((size_t)S.Np + 1, 0.0);
  // qt exists once a step has packed it; before that the divergence is of zero
  if (M.qt.size() != (size_t)S.Nqu + S.Nqv + S.Nqw + 1) M.qt.assign((size_t)S.Nqu + S.Nqv + S.Nqw + 1, 0.0);
  csr_spmv(S.D, M.qt.data(), M.divv.data());
  double d2 = 0.0;
  for (int i = 1; i <= S.Np; ++i) {
    double d = M.divv[(size_t)i] + M.ramp * S.div_bc[(size_t)i];
    d2 += d * d;
  }
  M.div_l2 = std::sqrt(d2 / std::max(1, S.Np));
  M.umax = M.vmax = M.wmax = 0.0;
  for (double x : M.u) M.umax = std::max(M.umax, std::abs(x));
  for (double x : M.v) M.vmax = std::max(M.vmax, std::abs(x));
  for (double x : M.w) M.wmax = std::max(M.wmax, std::abs(x));
  double fx = 0.0, fy = 0.0;
  for (auto &e : M.fu_w) fx += e.second * M.u[e.first];
  for (auto &e : M.px_w) fx += e.second * M.p[e.first];
  for (auto &e : M.fv_w) fy += e.second * M.v[e.first];
  for (auto &e : M.py_w) fy += e.second * M.p[e.first];
  const orc_params &p = S.par;
  double ub = 4.0 * p.Um / 9.0, den = p.rho * ub * ub * (2.0 * p.Rcirc) * p.Lz;
  M.fx = fx; M.fy = fy;
  M.cd = 2.0 * fx / den;
  M.cl = 2.0 * fy / den;
}

Sim *sim_create(const orc_params &par) {
  Sim *M = new Sim();
  M->S = setup_create(par);
  const Setup &S = *M->S;
  int nx = S.nx, ny = S.ny, nz = S.nz;
  M->Bu = box_from_aos(nx - 1, ny, nz, S.Au_vals.data());
  M->Bv = box_from_aos(nx, ny - 1, nz, S.Av_vals.data());
  M->Bw = box_from_aos(nx, ny, nz - 1, S.Aw_vals.data());
  M->Bp = box_from_aos(nx, ny, nz, S.Ap_vals.data());
  M->Mu.kind = M->Mv.kind = M->Mw.kind = 1;
  M->Mu.iters = M->Mv.iters = M->Mw.iters = std::max(1, par.vel_jacobi_iters);
  M->Mu.A = &M->Bu; M->Mv.A = &M->Bv; M->Mw.A = &M->Bw;
  M->Mp.kind = 2;
  M->Mp.iters = std::max(1, par.p_pc_maxit);
  M->Mp.A = &M->Bp;
  mg_setup(M->Mp.mg, M->Bp, par.p_pc_relax_type, 0);
  M->u.assign((size_t)(nx + 1) * ny * nz, 0.0);
  M->v.assign((size_t)nx * (ny + 1) * nz, 0.0);
  M->w.assign((size_t)nx * ny * (nz + 1), 0.0);
  M->p.assign((size_t)nx * ny * nz, 0.0);
  // Initial field: zero (reference, main.cu:123-126) or inlet profile on active u
  // faces plus seeded noise on active faces (production configuration).
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 1; i <= nx - 1; ++i) {
        size_t t = ixU(i, j, k, nx, ny);
        if (!S.active_u[t]) continue;
        double val = par.init_profile ? inlet_profile((j + 0.5) * S.dy, (k + 0.5) * S.dz, par.Ly, par.Um) : 0.0;
        if (par.noise_amp > 0.0) val += noise_at(par.seed, 0, t, par.noise_amp);
        M->u[t] = val;
      }
  if (par.noise_amp > 0.0) {
    for (int k = 0; k < nz; ++k)
      for (int j = 1; j <= ny - 1; ++j)
        for (int i = 0; i < nx; ++i) {
          size_t t = ixV(i, j, k, nx, ny);
          if (S.active_v[t]) M->v[t] = noise_at(par.seed, 1, t, par.noise_amp);
        }
    for (int k = 1; k <= nz - 1; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          size_t t = ixW(i, j, k, nx, ny);
          if (S.active_w[t]) M->w[t] = noise_at(par.seed, 2, t, par.noise_amp);
        }
  }
  // enforce_bc_and_cleanup with the unramped profile (immersed_operators.hpp:1036-1053)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) {
      M->u[ixU(0, j, k, nx, ny)] = S.uin[(size_t)j + (size_t)ny * k];
      M->u[ixU(nx, j, k, nx, ny)] = M->u[ixU(nx - 1, j, k, nx, ny)];
    }
  for (int k = 0; k < nz; ++k)
    for (int i = 0; i < nx; ++i) { M->v[ixV(i, 0, k, nx, ny)] = 0.0; M->v[ixV(i, ny, k, nx, ny)] = 0.0; }
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) { M->w[ixW(i, j, 0, nx, ny)] = 0.0; M->w[ixW(i, j, nz, nx, ny)] = 0.0; }
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 1; i <= nx - 1; ++i)
        if (!S.active_u[ixU(i, j, k, nx, ny)]) M->u[ixU(i, j, k, nx, ny)] = 0.0;
  for (int k = 0; k < nz; ++k)
    for (int j = 1; j <= ny - 1; ++j)
      for (int i = 0; i < nx; ++i)
        if (!S.active_v[ixV(i, j, k, nx, ny)]) M->v[ixV(i, j, k, nx, ny)] = 0.0;
  for (int k = 1; k <= nz - 1; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        if (!S.active_w[ixW(i, j, k, nx, ny)]) M->w[ixW(i, j, k, nx, ny)] = 0.0;
  M->uin_step = S.uin;
  force_weights(*M);
  return M;
}

int sim_phase(Sim &M, int ph) {
  switch (ph) {
    case 0: phase_ramp(M); return 0;
    case 1: phase_explicit(M); return 0;
    case 2: phase_momentum(M); return 0;
    case 3: phase_scatter(M); return 0;
    case 4: phase_divergence(M); return 0;
    case 5: phase_pressure(M); return 0;
    case 6: phase_projection(M); return 0;
    case 7: phase_diagnostics(M); return 0;
    default: return -1;
  }
}

// ---------------------------------------------------------------- named access
struct View {
  const void *p = nullptr;
  void *w = nullptr;
  long long n = -1;
  int es = 0;
};

template <class T>
View mk(const std::vector<T> &v) { View r; r.p = v.data(); r.n = (long long)v.size(); r.es = sizeof(T); return r; }
template <class T>
View mkw(std::vector<T> &v) { View r = mk(v); r.w = v.data(); return r; }

View setup_view(Setup &S, const std::string &n) {
  if (n == "beta2d") return mk(S.beta2d);
  if (n == "alpha_u2d") return mk(S.alpha_u2d);
  if (n == "alpha_v2d") return mk(S.alpha_v2d);
  if (n == "fluid2d") return mk(S.fluid2d);
  if (n == "fluid_p") return mk(S.fluid_p);
  if (n == "beta_p") return mk(S.beta_p);
  if (n == "active_u") return mk(S.active_u);
  if (n == "active_v") return mk(S.active_v);
  if (n == "active_w") return mk(S.active_w);
  if (n == "active_qu") return mk(S.active_qu);
  if (n == "active_qv") return mk(S.active_qv);
  if (n == "active_qw") return mk(S.active_qw);
  if (n == "id_u") return mk(S.id_u);
  if (n == "id_v") return mk(S.id_v);
  if (n == "id_w") return mk(S.id_w);
  if (n == "id_p") return mk(S.id_p);
  if (n == "id_qu") return mk(S.id_qu);
  if (n == "id_qv") return mk(S.id_qv);
  if (n == "id_qw") return mk(S.id_qw);
  if (n == "rhs_u_bc") return mk(S.rhs_u_bc);
  if (n == "rhs_v_bc") return mk(S.rhs_v_bc);
  if (n == "rhs_w_bc") return mk(S.rhs_w_bc);
  if (n == "div_bc") return mk(S.div_bc);
  if (n == "scale_u") return mk(S.scale_u);
  if (n == "scale_v") return mk(S.scale_v);
  if (n == "scale_w") return mk(S.scale_w);
  if (n == "Au_vals") return mk(S.Au_vals);
  if (n == "Av_vals") return mk(S.Av_vals);
  if (n == "Aw_vals") return mk(S.Aw_vals);
  if (n == "Ap_vals") return mk(S.Ap_vals);
  if (n == "D_rowptr") return mk(S.D.rowptr);
  if (n == "D_colind") return mk(S.D.colind);
  if (n == "D_vals") return mk(S.D.vals);
  if (n == "G_rowptr") return mk(S.G.rowptr);
  if (n == "G_colind") return mk(S.G.colind);
  if (n == "G_vals") return mk(S.G.vals);
  if (n == "uin") return mk(S.uin);
  return View();
}

View sim_view(Sim &M, const std::string &n) {
  if (n == "u") return mkw(M.u);
  if (n == "v") return mkw(M.v);
  if (n == "w") return mkw(M.w);
  if (n == "p") return mkw(M.p);
  if (n == "u_tilde") return mkw(M.ut);
  if (n == "v_tilde") return mkw(M.vt);
  if (n == "w_tilde") return mkw(M.wt);
  if (n == "conv_u") return mkw(M.conv_u);
  if (n == "conv_v") return mkw(M.conv_v);
  if (n == "conv_w") return mkw(M.conv_w);
  if (n == "gpx") return mkw(M.gpx);
  if (n == "gpy") return mkw(M.gpy);
  if (n == "gpz") return mkw(M.gpz);
  if (n == "rhs_u") return mkw(M.rhs_u);
  if (n == "rhs_v") return mkw(M.rhs_v);
  if (n == "rhs_w") return mkw(M.rhs_w);
  if (n == "rhs_p") return mkw(M.rhs_p);
  if (n == "x_u") return mkw(M.x_u);
  if (n == "x_v") return mkw(M.x_v);
  if (n == "x_w") return mkw(M.x_w);
  if (n == "x_p") return mkw(M.x_p);
  if (n == "qt") return mkw(M.qt);
  if (n == "Dqt") return mkw(M.Dqt);
  if (n == "divv") return mkw(M.divv);
  if (n == "uin_step") return mkw(M.uin_step);
  return setup_view(*M.S, n);
}

int copy_out(const View &v, void *dst, long long nbytes) {
  if (v.n < 0 || !v.p) return -1;
  long long need = v.n * v.es;
  if (nbytes < need) return -2;
  std::memcpy(dst, v.p, (size_t)need);
  return 0;
}

}  // namespace

// ---------------------------------------------------------------- C API
extern "C" {

void orc_set_coarse_size(int n) { g_coarse = n > 0 ? (size_t)n : 0; }

void orc_set_threads(int n) {
  g_threads = n < 1 ? 1 : n;
  omp_set_dynamic(0);
}

void orc_default_params(orc_params *p) {
  // common.hpp:45-85
  std::memset(p, 0, sizeof(*p));
  p->nx = 32; p->ny = 12; p->nz = 12;
  p->Lx = 2.2; p->Ly = 0.41; p->Lz = 0.41;
  p->rho = 1.0; p->nu = 1.0e-3; p->Um = 0.45; p->dt = 2.0e-3;
  p->nsteps = 10; p->print_every = 1; p->inlet_ramp_steps = 200; p->use_zero_initial_guess = 1;
  p->nsub_beta = 10; p->beta_tol = 1.0e-8; p->alpha_tol = 1.0e-10;
  p->xcirc = 0.2; p->ycirc = 0.2; p->Rcirc = 0.05;
  p->vel_maxit = 250; p->vel_kdim = 50; p->vel_tol = 1.0e-6; p->vel_jacobi_iters = 1;
  p->p_maxit = 250; p->p_kdim = 50; p->p_tol = 1.0e-9;
  p->p_pc_maxit = 1; p->p_pc_relax_type = 1; p->p_pc_rap_type = 1;
  p->init_profile = 0; p->noise_amp = 0.0; p->seed = 0;
}

void *orc_setup_create(const orc_params *par) { return setup_create(*par); }
void orc_setup_destroy(void *h) { delete static_cast<Setup *>(h); }

long long orc_setup_len(void *h, const char *name) { return setup_view(*static_cast<Setup *>(h), name).n; }
int orc_setup_elsize(void *h, const char *name) { return setup_view(*static_cast<Setup *>(h), name).es; }
int orc_setup_copy(void *h, const char *name, void *dst, long long nbytes) {
  return copy_out(setup_view(*static_cast<Setup *>(h), name), dst, nbytes);
}
long long orc_setup_int(void *h, const char *name) {
  Setup &S = *static_cast<Setup *>(h);
  std::string n(name);
  if (n == "Nu") return S.Nu;
  if (n == "Nv") return S.Nv;
  if (n == "Nw") return S.Nw;
  if (n == "Np") return S.Np;
  if (n == "Nqu") return S.Nqu;
  if (n == "Nqv") return S.Nqv;
  if (n == "Nqw") return S.Nqw;
  if (n == "n_pruned") return S.n_pruned;
  return -1;
}

void orc_struct_matvec(int nx, int ny, int nz, const double *vals, const double *x, double *y) {
  Box B = box_from_aos(nx, ny, nz, vals);
  box_matvec(B, x, y);
}

void orc_csr_spmv_1based(int nrows, const int *rowptr, const int *colind, const double *vals, const double *x, double *y) {
  for (int r = 0; r < nrows; ++r) {
    double s = 0.0;
    for (int jj = rowptr[r]; jj < rowptr[r + 1]; ++jj) s += vals[jj] * x[colind[jj]];
    y[r + 1] = s;
  }
}

int orc_gmres(int nx, int ny, int nz, const double *vals, const double *b, double *x, double tol, int maxit, int kdim,
              int precond, int pc_iters, int relax_type, int rap_type, int *its, double *relres) {
  (void)rap_type;
  Box B = box_from_aos(nx, ny, nz, vals);
  Precond M;
  M.kind = precond;
  M.iters = std::max(1, pc_iters);
  M.A = &B;
  if (precond == 2) mg_setup(M.mg, B, relax_type, 0);
  return gmres_solve(B, M, b, x, tol, maxit, kdim, its, relres);
}

int orc_precond_apply(int nx, int ny, int nz, const double *vals, const double *b, double *x, int precond, int pc_iters,
                      int relax_type, int rap_type) {
  (void)rap_type;
  Box B = box_from_aos(nx, ny, nz, vals);
  Precond M;
  M.kind = precond;
  M.iters = std::max(1, pc_iters);
  M.A = &B;
  if (precond == 2) mg_setup(M.mg, B, relax_type, 0);
  M.apply(b, x);
  return 0;
}

int orc_mg_levels(int nx, int ny, int nz, const double *vals, int rap_type, int max_levels, int *dims, int *cdir, double *omega) {
  (void)rap_type;
  Box B = box_from_aos(nx, ny, nz, vals);
  MG M;
  mg_setup(M, B, 1, max_levels);
  int nl = (int)M.L.size();
  for (int l = 0; l < nl; ++l) {
    dims[3 * l] = M.L[(size_t)l].A.nx;
    dims[3 * l + 1] = M.L[(size_t)l].A.ny;
    dims[3 * l + 2] = M.L[(size_t)l].A.nz;
    cdir[l] = M.L[(size_t)l].cdir;
    omega[l] = M.L[(size_t)l].omega;
  }
  return nl;
}

int orc_mg_level_matrix(int nx, int ny, int nz, const double *vals, int relax_type, int level, double *out) {
  Box B = box_from_aos(nx, ny, nz, vals);
  MG M;
  mg_setup(M, B, relax_type, 0);
  if (level < 0 || level >= (int)M.L.size()) return -1;
  const Box &L = M.L[(size_t)level].A;
  for (size_t r = 0; r < L.n(); ++r)
    for (int s = 0; s < 7; ++s) out[7 * r + s] = L.a[s][r];
  return 0;
}

void *orc_sim_create(const orc_params *par) { return sim_create(*par); }
void orc_sim_destroy(void *s) {
  Sim *M = static_cast<Sim *>(s);
  delete M->S;
  delete M;
}
int orc_sim_step(void *s, int nsteps) {
  Sim &M = *static_cast<Sim *>(s);
  for (int n = 0; n < nsteps; ++n)
    for (int ph = 0; ph < 8; ++ph) sim_phase(M, ph);
  return 0;
}
int orc_sim_phase(void *s, int phase) { return sim_phase(*static_cast<Sim *>(s), phase); }
long long orc_sim_len(void *s, const char *name) { return sim_view(*static_cast<Sim *>(s), name).n; }
int orc_sim_copy(void *s, const char *name, void *dst, long long nbytes) {
  return copy_out(sim_view(*static_cast<Sim *>(s), name), dst, nbytes);
}
int orc_sim_set(void *s, const char *name, const void *src, long long nbytes) {
  if (std::string(name) == "inlet_amplitude") {
    // theta (ny*nz): the inlet velocity and the divergence inlet term become theta times
    // their reference values (the adjoint's parameter; no reference counterpart)
    Sim &M = *static_cast<Sim *>(s);
    Setup &S = *M.S;
    const size_t n = (size_t)S.ny * S.nz;
    if ((long long)(n * sizeof(double)) != nbytes) return -1;
    if (M.uin_base.empty()) {
      M.uin_base = S.uin;
      M.div_bc_base = S.div_bc;
    }
    const double *th = static_cast<const double *>(src);
    for (size_t q = 0; q < n; ++q) S.uin[q] = M.uin_base[q] * th[q];
    for (int k = 0; k < S.nz; ++k)
      for (int j = 0; j < S.ny; ++j) {
        size_t c = ixP(0, j, k, S.nx, S.ny);
        if (!S.fluid_p[c]) continue;
        size_t row = (size_t)S.id_p[c];
        S.div_bc[row] = M.div_bc_base[row] * th[(size_t)j + (size_t)S.ny * k];
      }
    return 0;
  }
  View v = sim_view(*static_cast<Sim *>(s), name);
  if (!v.w || v.n * v.es != nbytes) return -1;
  std::memcpy(v.w, src, (size_t)nbytes);
  return 0;
}
// Cell-centred ASCII VTK file (immersed_operators.hpp:1055-1076): fprintf of every
// value with "%.16e", face averages for ux, uy, uz, then pressure and the fluid mask.
int orc_write_vtk_cellcentered(const char *filename, const double *u, const double *v, const double *w,
                               const double *p, const unsigned char *fluid_p, int nx, int ny, int nz, double Lx,
                               double Ly, double Lz) {
  double dx = Lx / nx, dy = Ly / ny, dz = Lz / nz;
  std::FILE *fp = std::fopen(filename, "w");
  if (!fp) return 1;
  std::fprintf(fp, "# vtk DataFile Version 3.0\n");
  std::fprintf(fp, "3D immersed cylinder MAC cell-centered\nASCII\nDATASET STRUCTURED_POINTS\n");
  std::fprintf(fp, "DIMENSIONS %d %d %d\n", nx + 1, ny + 1, nz + 1);
  std::fprintf(fp, "ORIGIN 0 0 0\nSPACING %.16e %.16e %.16e\n", dx, dy, dz);
  std::fprintf(fp, "CELL_DATA %d\n", nx * ny * nz);
  std::fprintf(fp, "SCALARS ux double 1\nLOOKUP_TABLE default\n");
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) std::fprintf(fp, "%.16e\n", 0.5 * (u[ixU(i, j, k, nx, ny)] + u[ixU(i + 1, j, k, nx, ny)]));
  std::fprintf(fp, "SCALARS uy double 1\nLOOKUP_TABLE default\n");
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) std::fprintf(fp, "%.16e\n", 0.5 * (v[ixV(i, j, k, nx, ny)] + v[ixV(i, j + 1, k, nx, ny)]));
  std::fprintf(fp, "SCALARS uz double 1\nLOOKUP_TABLE default\n");
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) std::fprintf(fp, "%.16e\n", 0.5 * (w[ixW(i, j, k, nx, ny)] + w[ixW(i, j, k + 1, nx, ny)]));
  std::fprintf(fp, "SCALARS pressure double 1\nLOOKUP_TABLE default\n");
  for (size_t c = 0; c < (size_t)nx * ny * nz; ++c) std::fprintf(fp, "%.16e\n", p[c]);
  std::fprintf(fp, "SCALARS fluid_mask int 1\nLOOKUP_TABLE default\n");
  for (size_t c = 0; c < (size_t)nx * ny * nz; ++c) std::fprintf(fp, "%d\n", fluid_p[c] ? 1 : 0);
  std::fclose(fp);
  return 0;
}

double orc_sim_scalar(void *s, const char *name) {
  Sim &M = *static_cast<Sim *>(s);
  std::string n(name);
  if (n == "ramp") return M.ramp;
  if (n == "div_l2") return M.div_l2;
  if (n == "umax") return M.umax;
  if (n == "vmax") return M.vmax;
  if (n == "wmax") return M.wmax;
  if (n == "relres_u") return M.rel[0];
  if (n == "relres_v") return M.rel[1];
  if (n == "relres_w") return M.rel[2];
  if (n == "relres_p") return M.rel[3];
  if (n == "its_u") return M.its[0];
  if (n == "its_v") return M.its[1];
  if (n == "its_w") return M.its[2];
  if (n == "its_p") return M.its[3];
  if (n == "err_u") return M.err[0];
  if (n == "err_v") return M.err[1];
  if (n == "err_w") return M.err[2];
  if (n == "err_p") return M.err[3];
  if (n == "step") return M.step;
  if (n == "fx") return M.fx;
  if (n == "fy") return M.fy;
  if (n == "cd") return M.cd;
  if (n == "cl") return M.cl;
  return NAN;
}

}  // extern "C"

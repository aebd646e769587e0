// pfmg.hip -- PFMG-style semicoarsening multigrid (libfpc).
//
// Replaces HYPRE_StructPFMG* as the pressure preconditioner of the reference
// (main.cu:213-220: MaxIter 1, Tol 0, zero guess, RelaxType, RAPType).  HYPRE's
// PFMG internals are not part of the reference and its version is unpinned, so
// this is a PFMG-style method with the build's own, fully stated definition:
//
//  * each level halves one direction d: the one with the largest total in-box
//    coupling c_d = sum over rows (k, j, i order) of |a_-d| + |a_+d|; a later
//    direction must exceed the best so far by a relative 1e-9 to win.  Coarse
//    point c sits on fine index 2c+1, nc = floor(n/2);
//  * interpolation at an F point (even index along d): w-/+ = -a_-/+d / centre,
//    centre = diagonal + the couplings across d (collapsed); C points inject;
//    restriction is the transpose;
//  * coarse operator (non-Galerkin, 7-point on every level, RAP type 0 and 1):
//    couplings across d are the Galerkin couplings with that cross offset summed
//    over the offsets along d; couplings along d are the 1-D Galerkin product of
//    the collapsed stencil; the centre keeps the Galerkin row sum.  For the
//    7-point Laplacian this is twice the operator rediscretised on the
//    semicoarsened grid, which keeps the V-cycle rate mesh-independent;
//  * relaxation: Jacobi (type 0), weighted Jacobi with w = 2/(2 + c_d/sum c)
//    (type 1), red/black Gauss-Seidel symmetric (2) or forward (3);
//  * coarsening stops once a level has at most min(1152, max(64, N0/2048))
//    points (N0: the finest level's), or when no direction has 2+ points; that
//    level is solved exactly with a dense inverse held in FP64 (one GEMV).
//    Large boxes thus skip the deepest, latency-bound levels (512x96x96: 13
//    levels instead of 18, same GMRES iteration count); small boxes keep the
//    deep hierarchy they need.
//
// The hierarchy is built once per matrix on the host in FP64 and uploaded as
// x-contiguous planes; in FP32 mode every level's coefficients, smoothing,
// restriction and prolongation run in FP32 (half the bytes of this
// bandwidth-bound cycle) while the coarsest solve stays FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "extrude.hpp"
#include "fpc_internal.hpp"
#include "batched.hpp"
#include "mg_device.hpp"
#include "stencil_tile.hpp"

namespace fpc {
namespace {

// Largest coarsest level for a hierarchy whose finest level has n0 points.
inline long long coarse_max(long long n0) { return std::min(1152LL, std::max(64LL, n0 / 2048)); }
// Upper bound of a coarsest size set by the caller (dense inverse: n^2 doubles).
constexpr long long kCoarseSizeLimit = 4096;

inline int em(int d) { return 1 + 2 * d; }
inline int ep(int d) { return 2 + 2 * d; }

// ---------------------------------------------------------------- host hierarchy
struct HLevel {
  int n[3] = {0, 0, 0};
  long long N = 0;
  std::vector<double> a;  // 7 * N, slot-major planes
  int cdir = -1;
  double omega = 1.0;
  std::vector<double> wm, wp;
  double A(int s, long long r) const { return a[(size_t)s * (size_t)N + (size_t)r]; }
};

inline long long lin(const HLevel &L, int i, int j, int k) { return ((long long)k * L.n[1] + j) * L.n[0] + i; }

// Coefficient of slot s at point p, zero when the neighbour leaves the box.
inline double coef_inbox(const HLevel &L, const int *p, int s) {
  if (s != 0) {
    int d = (s - 1) / 2, sg = (s & 1) ? -1 : 1;
    int q = p[d] + sg;
    if (q < 0 || q >= L.n[d]) return 0.0;
  }
  return L.A(s, lin(L, p[0], p[1], p[2]));
}

void choose_direction(HLevel &L) {
  double c[3] = {0.0, 0.0, 0.0};
  int p[3];
  for (p[2] = 0; p[2] < L.n[2]; ++p[2])
    for (p[1] = 0; p[1] < L.n[1]; ++p[1])
      for (p[0] = 0; p[0] < L.n[0]; ++p[0])
        for (int d = 0; d < 3; ++d) c[d] += std::abs(coef_inbox(L, p, em(d))) + std::abs(coef_inbox(L, p, ep(d)));
  L.cdir = -1;
  for (int d = 0; d < 3; ++d) {
    if (L.n[d] < 2) continue;
    if (L.cdir < 0 || c[d] > c[L.cdir] * (1.0 + 1.0e-9)) L.cdir = d;
  }
  double tot = c[0] + c[1] + c[2];
  double s = (L.cdir >= 0 && tot > 0.0) ? c[L.cdir] / tot : 1.0;
  L.omega = 2.0 / (2.0 + s);
}

void interp_weights(HLevel &L) {
  int d = L.cdir;
  L.wm.assign((size_t)L.N, 0.0);
  L.wp.assign((size_t)L.N, 0.0);
  int p[3];
  for (p[2] = 0; p[2] < L.n[2]; ++p[2])
    for (p[1] = 0; p[1] < L.n[1]; ++p[1])
      for (p[0] = 0; p[0] < L.n[0]; ++p[0]) {
        if (p[d] & 1) continue;
        long long r = lin(L, p[0], p[1], p[2]);
        double ctr = L.A(0, r);
        for (int s = 1; s < kSlots; ++s)
          if (s != em(d) && s != ep(d)) ctr += coef_inbox(L, p, s);
        if (ctr != 0.0) {
          L.wm[(size_t)r] = -coef_inbox(L, p, em(d)) / ctr;
          L.wp[(size_t)r] = -coef_inbox(L, p, ep(d)) / ctr;
        }
      }
}

HLevel coarsen(const HLevel &F) {
  int d = F.cdir;
  HLevel C;
  for (int q = 0; q < 3; ++q) C.n[q] = F.n[q];
  C.n[d] = F.n[d] / 2;
  C.N = (long long)C.n[0] * C.n[1] * C.n[2];
  C.a.assign((size_t)kSlots * (size_t)C.N, 0.0);
  int cp[3];
  for (cp[2] = 0; cp[2] < C.n[2]; ++cp[2])
    for (cp[1] = 0; cp[1] < C.n[1]; ++cp[1])
      for (cp[0] = 0; cp[0] < C.n[0]; ++cp[0]) {
        int f0[3] = {cp[0], cp[1], cp[2]};
        f0[d] = 2 * cp[d] + 1;
        int fm[3] = {f0[0], f0[1], f0[2]}, fp[3] = {f0[0], f0[1], f0[2]};
        fm[d] -= 1;
        fp[d] += 1;
        bool hasp = fp[d] < F.n[d];
        long long rfm = lin(F, fm[0], fm[1], fm[2]);
        long long rfp = hasp ? lin(F, fp[0], fp[1], fp[2]) : -1;
        double rm = F.wp[(size_t)rfm];
        double rp = hasp ? F.wm[(size_t)rfp] : 0.0;
        double a0m = coef_inbox(F, fm, 0), a0p = hasp ? coef_inbox(F, fp, 0) : 0.0;
        double ac = coef_inbox(F, f0, 0) + rm * (coef_inbox(F, f0, em(d)) + coef_inbox(F, fm, ep(d)) + rm * a0m);
        if (hasp) ac += rp * (coef_inbox(F, f0, ep(d)) + coef_inbox(F, fp, em(d)) + rp * a0p);
        // along d: Galerkin form (for the row sum) and the 1-D Galerkin product of
        // the collapsed stencil (what is stored)
        double wmm = F.wm[(size_t)rfm];
        double amG = coef_inbox(F, f0, em(d)) * wmm + rm * a0m * wmm + rm * coef_inbox(F, fm, em(d));
        double cmm = a0m, cpp = a0p;
        for (int s = 1; s < kSlots; ++s) {
          if (s == em(d) || s == ep(d)) continue;
          cmm += coef_inbox(F, fm, s);
          if (hasp) cpp += coef_inbox(F, fp, s);
        }
        double am = coef_inbox(F, f0, em(d)) * wmm + rm * cmm * wmm + rm * coef_inbox(F, fm, em(d));
        double apG = 0.0, apv = 0.0;
        if (hasp) {
          double wpp = F.wp[(size_t)rfp];
          apG = coef_inbox(F, f0, ep(d)) * wpp + rp * a0p * wpp + rp * coef_inbox(F, fp, ep(d));
          apv = coef_inbox(F, f0, ep(d)) * wpp + rp * cpp * wpp + rp * coef_inbox(F, fp, ep(d));
        }
        long long rc = lin(C, cp[0], cp[1], cp[2]);
        // across d: Galerkin couplings with this cross offset, summed over the offsets along d
        for (int s = 1; s < kSlots; ++s) {
          if (s == em(d) || s == ep(d)) continue;
          int e = (s - 1) / 2, sg = (s & 1) ? -1 : 1;
          double v = coef_inbox(F, f0, s);
          double cm = coef_inbox(F, fm, s);
          if (cm != 0.0) {
            int q[3] = {fm[0], fm[1], fm[2]};
            q[e] += sg;
            long long rq = lin(F, q[0], q[1], q[2]);
            v += rm * cm * (F.wm[(size_t)rq] + F.wp[(size_t)rq]);
          }
          if (hasp) {
            double cq = coef_inbox(F, fp, s);
            if (cq != 0.0) {
              int q[3] = {fp[0], fp[1], fp[2]};
              q[e] += sg;
              long long rq = lin(F, q[0], q[1], q[2]);
              v += rp * cq * (F.wm[(size_t)rq] + F.wp[(size_t)rq]);
            }
          }
          C.a[(size_t)s * (size_t)C.N + (size_t)rc] = v;
        }
        // centre: keeps the Galerkin row sum
        C.a[(size_t)rc] = ac + (amG - am) + (apG - apv);
        C.a[(size_t)em(d) * (size_t)C.N + (size_t)rc] = am;
        C.a[(size_t)ep(d) * (size_t)C.N + (size_t)rc] = apv;
      }
  return C;
}

// Dense inverse of the coarsest operator: PA = LU with partial pivoting, the
// elimination confined to the band that pivoting can reach (bw below, 2 bw
// above the diagonal; bw = nx*ny for a 7-point box in x-fastest order), then
// the unit columns solved in parallel host threads.
std::vector<double> dense_inverse(const HLevel &L) {
  const int n = (int)L.N;
  const int bw = (L.n[2] > 1) ? L.n[0] * L.n[1] : (L.n[1] > 1 ? L.n[0] : 1);
  std::vector<double> M((size_t)n * n, 0.0), Inv((size_t)n * n, 0.0);
  std::vector<int> piv((size_t)n, 0);
  int p[3];
  for (p[2] = 0; p[2] < L.n[2]; ++p[2])
    for (p[1] = 0; p[1] < L.n[1]; ++p[1])
      for (p[0] = 0; p[0] < L.n[0]; ++p[0]) {
        long long r = lin(L, p[0], p[1], p[2]);
        M[(size_t)r * n + r] += L.A(0, r);
        for (int s = 1; s < kSlots; ++s) {
          int d = (s - 1) / 2, sg = (s & 1) ? -1 : 1;
          int q[3] = {p[0], p[1], p[2]};
          q[d] += sg;
          if (q[d] < 0 || q[d] >= L.n[d]) continue;
          M[(size_t)r * n + lin(L, q[0], q[1], q[2])] += L.A(s, r);
        }
      }
  for (int c = 0; c < n; ++c) {
    const int rmax = std::min(n - 1, c + bw), tmax = std::min(n - 1, c + 2 * bw);
    int pv = c;
    for (int r = c + 1; r <= rmax; ++r)
      if (std::abs(M[(size_t)r * n + c]) > std::abs(M[(size_t)pv * n + c])) pv = r;
    piv[(size_t)c] = pv;
    if (pv != c)
      for (int t = 0; t < n; ++t) std::swap(M[(size_t)c * n + t], M[(size_t)pv * n + t]);
    const double dg = M[(size_t)c * n + c];
    if (dg == 0.0) continue;
    for (int r = c + 1; r <= rmax; ++r) {
      const double f = M[(size_t)r * n + c] / dg;
      M[(size_t)r * n + c] = f;
      if (f == 0.0) continue;
      for (int t = c + 1; t <= tmax; ++t) M[(size_t)r * n + t] -= f * M[(size_t)c * n + t];
    }
  }
  auto solve_columns = [&](int j0, int j1) {
    std::vector<double> x((size_t)n);
    for (int j = j0; j < j1; ++j) {
      std::fill(x.begin(), x.end(), 0.0);
      x[(size_t)j] = 1.0;
      for (int c = 0; c < n; ++c) {
        const int pv = piv[(size_t)c];
        if (pv != c) std::swap(x[(size_t)c], x[(size_t)pv]);
        const double xc = x[(size_t)c];
        if (xc == 0.0) continue;
        for (int r = c + 1; r < n; ++r) x[(size_t)r] -= M[(size_t)r * n + c] * xc;
      }
      for (int r = n - 1; r >= 0; --r) {
        double sum = x[(size_t)r];
        for (int t = r + 1; t < n; ++t) sum -= M[(size_t)r * n + t] * x[(size_t)t];
        const double dg = M[(size_t)r * n + r];
        x[(size_t)r] = (dg != 0.0) ? sum / dg : 0.0;
      }
      for (int r = 0; r < n; ++r) Inv[(size_t)r * n + j] = x[(size_t)r];
    }
  };
  const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 256 || nth == 1) {
    solve_columns(0, n);
  } else {
    std::vector<std::thread> pool;
    for (int q = 0; q < nth; ++q) pool.emplace_back(solve_columns, (int)((long long)n * q / nth),
                                                    (int)((long long)n * (q + 1) / nth));
    for (auto &t : pool) t.join();
  }
  return Inv;
}

// ---------------------------------------------------------------- device kernels
// Read when a kernel gets no skip flag (always 0).
__device__ int g_no_skip = 0;

// The one-point kernels of the small levels are latency-bound: they read the
// skip flag with a vector load issued together with their data loads and test
// it only before the store, so the flag costs no round trip of its own.
__device__ inline int skip_flag(const int *skip) {
  return __hip_atomic_load(skip ? skip : &g_no_skip, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
__global__ void k_relax_zero(const T *__restrict__ dinv, const T *__restrict__ b, T *__restrict__ x, T w, long long n,
                             const int *skip) {
  if (skip && *skip) return;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
    x[t] = w * b[t] * dinv[t];
}

template <typename T>
__device__ inline T apply7(const T *__restrict__ c, long long pitch, long long r, int i, int j, int k, int nx, int ny,
                           int nz, const T *__restrict__ x) {
  long long pl = (long long)nx * ny;
  T s = c[r] * x[r];
  if (i > 0) s += c[pitch + r] * x[r - 1];
  if (i < nx - 1) s += c[2 * pitch + r] * x[r + 1];
  if (j > 0) s += c[3 * pitch + r] * x[r - nx];
  if (j < ny - 1) s += c[4 * pitch + r] * x[r + nx];
  if (k > 0) s += c[5 * pitch + r] * x[r - pl];
  if (k < nz - 1) s += c[6 * pitch + r] * x[r + pl];
  return s;
}

template <typename T>
__global__ void k_residual_mg(const T *__restrict__ c, long long pitch, int nx, int ny, int nz, const T *__restrict__ b,
                              const T *__restrict__ x, T *__restrict__ r, const int *skip) {
  if (skip && *skip) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  if (i >= nx || j >= ny) return;
  long long t = ((long long)k * ny + j) * nx + i;
  r[t] = b[t] - apply7(c, pitch, t, i, j, k, nx, ny, nz, x);
}

template <typename T>
__global__ void k_relax_jacobi(const T *__restrict__ c, long long pitch, int nx, int ny, int nz,
                               const T *__restrict__ dinv, const T *__restrict__ b, const T *__restrict__ x,
                               T *__restrict__ y, T w, const int *skip) {
  if (skip && *skip) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  if (i >= nx || j >= ny) return;
  long long t = ((long long)k * ny + j) * nx + i;
  y[t] = x[t] + w * (b[t] - apply7(c, pitch, t, i, j, k, nx, ny, nz, x)) * dinv[t];
}

// Red/black sweep of one colour, in place.  first: the other colour is zero and
// this colour starts from zero (zero-guess first half-sweep).
template <typename T>
__global__ void k_relax_rb(const T *__restrict__ c, long long pitch, int nx, int ny, int nz, const T *__restrict__ dinv,
                           const T *__restrict__ b, T *__restrict__ x, int color, int first, const int *skip) {
  if (skip && *skip) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  if (i >= nx || j >= ny) return;
  long long t = ((long long)k * ny + j) * nx + i, pl = (long long)nx * ny;
  if (((i + j + k) & 1) != color) {
    if (first) x[t] = (T)0;
    return;
  }
  if (first) {
    x[t] = b[t] * dinv[t];
    return;
  }
  T s = b[t];
  if (i > 0) s -= c[pitch + t] * x[t - 1];
  if (i < nx - 1) s -= c[2 * pitch + t] * x[t + 1];
  if (j > 0) s -= c[3 * pitch + t] * x[t - nx];
  if (j < ny - 1) s -= c[4 * pitch + t] * x[t + nx];
  if (k > 0) s -= c[5 * pitch + t] * x[t - pl];
  if (k < nz - 1) s -= c[6 * pitch + t] * x[t + pl];
  x[t] = s * dinv[t];
}

// b_c = r(F0) + w+(Fm) r(Fm) + w-(Fp) r(Fp), one thread per coarse point.
template <typename T>
__global__ void k_restrict(int fnx, int fny, int fnz, int d, const T *__restrict__ r, const T *__restrict__ wm,
                           const T *__restrict__ wp, int cnx, int cny, int cnz, T *__restrict__ bc, const int *skip) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  if (i >= cnx || j >= cny) return;
  const int sk = skip_flag(skip);
  const T v = mgd::restrict_point<T>(fnx, fny, fnz, d, r, wm, wp, FullMap(), i, j, k);
  pin(v);
  if (sk) return;
  bc[((long long)k * cny + j) * cnx + i] = v;
}

// x_f += P x_c, one thread per fine point.
template <typename T>
__global__ void k_prolong_add(int fnx, int fny, int fnz, int d, const T *__restrict__ wm, const T *__restrict__ wp,
                              const T *__restrict__ xc, int cnx, int cny, int cnz, T *__restrict__ xf,
                              const int *skip) {
  if (skip && *skip) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  if (i >= fnx || j >= fny) return;
  int p[3] = {i, j, k};
  int nc[3] = {cnx, cny, cnz};
  long long t = ((long long)k * fny + j) * fnx + i;
  long long cstride = (d == 0) ? 1 : (d == 1 ? (long long)cnx : (long long)cnx * cny);
  int q[3] = {i, j, k};
  if (p[d] & 1) {
    q[d] = (p[d] - 1) >> 1;
    xf[t] += xc[((long long)q[2] * cny + q[1]) * cnx + q[0]];
  } else {
    q[d] = p[d] >> 1;  // right coarse neighbour
    long long cr = ((long long)q[2] * cny + q[1]) * cnx + q[0];
    T s = (T)0;
    if (q[d] - 1 >= 0) s += wm[t] * xc[cr - cstride];
    if (q[d] < nc[d]) s += wp[t] * xc[cr];
    xf[t] += s;
  }
}

// Down-sweep and restriction of a small level in one launch, one thread per
// coarse point (fine box nx x ny x nz, coarse cnx x cny x cnz).
// TB: storage type of b (wider than T on the finest level of a reduced-precision cycle).
template <typename T, typename TB = T>
__global__ void __launch_bounds__(256) k_down_restrict(const T *__restrict__ c, long long pitch, int nx, int ny, int nz,
                                const T *__restrict__ dinv, const TB *__restrict__ b, const T *__restrict__ wm,
                                const T *__restrict__ wp, int d, T w, int cnx, int cny, T *__restrict__ bc,
                                const int *skip) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  if (i >= cnx || j >= cny) return;
  const int sk = skip_flag(skip);
  const T v = mgd::down_restrict_point<T, FullMap, TB>(c, pitch, FullMap(), nx, ny, nz, dinv, b, w, wm, wp, d, i, j, k);
  pin(v);
  if (sk) return;
  bc[((long long)k * cny + j) * cnx + i] = v;
}

// TB, TO: storage types of b and y (wider than T on the finest level of a
// reduced-precision cycle, which then reads and writes the FP64 vectors directly).
template <typename T, int D, typename TB = T, typename TO = T>
__global__ void __launch_bounds__(256) k_up_jacobi(const T *__restrict__ c, long long pitch, int nx, int ny, int nz,
                            const T *__restrict__ dinv, const TB *__restrict__ b, const T *__restrict__ wm,
                            const T *__restrict__ wp, const T *__restrict__ yc, int cnx, int cny, int ncd,
                            TO *__restrict__ y, T w, const int *skip) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  if (i >= nx || j >= ny) return;
  long long t = ((long long)k * ny + j) * nx + i;
  const int sk = skip_flag(skip);
  const T v =
      mgd::up_point<T, D, FullMap, TB>(c, pitch, FullMap(), nx, ny, nz, dinv, b, wm, wp, yc, cnx, cny, ncd, w, i, j, k, t);
  pin(v);
  if (sk) return;
  y[t] = (TO)v;
}

// x = Ainv b on the coarsest level (FP64 inverse), one block.
// ---------------------------------------------------------------- z-marching forms
// The same arithmetic as jacobi_resid / k_restrict / k_up_jacobi, on the
// z-marching tiles: every row read once, z-neighbours carried in registers.

template <typename T, int TY, typename TB = T>
__global__ void __launch_bounds__(kTileX *TY)
    k_down_tiled(const T *__restrict__ c, long long pitch, ZMap zm, int nx, int ny, int nz, const T *__restrict__ dinv,
                 const TB *__restrict__ b, T *__restrict__ r, T w, int kz, const int *skip) {
  if (skip && *skip) return;
  mgd::DownOp<T, TB> op{c, pitch, nx, ny, dinv, b, r, w, zm};
  const int k0 = blockIdx.z * kz;
  march_tile_pipelined<T, kTileX, TY>(op, nx, ny, nz, k0, min(k0 + kz, nz), blockIdx.x, blockIdx.y);
}

template <typename T, int TY, typename TB = T>
__global__ void __launch_bounds__(kTileX *TY)
    k_down_restrict_z(const T *__restrict__ c, long long pitch, ZMap zm, int nx, int ny, int nz,
                      const T *__restrict__ dinv, const TB *__restrict__ b, const T *__restrict__ wm,
                      const T *__restrict__ wp, T *__restrict__ bc, int ncz, T w, int kz, const int *skip) {
  if (skip && *skip) return;
  mgd::DownRestrictZOp<T, TB> op;
  op.c = c; op.pitch = pitch; op.nx = nx; op.ny = ny; op.dinv = dinv; op.b = b; op.w = w;
  op.zm = zm;
  op.wm = wm; op.wp = wp; op.bc = bc; op.nz = nz; op.ncz = ncz;
  const int k0 = blockIdx.z * kz, k1 = min(k0 + kz, nz);
  op.k0 = k0;
  march_tile_pipelined<T, kTileX, TY>(op, nx, ny, nz, k0, k1 < nz ? k1 + 1 : k1, blockIdx.x, blockIdx.y);
}

template <typename T, int D, int TY, typename TB = T, typename TO = T>
__global__ void __launch_bounds__(kTileX *TY)
    k_up_tiled(const T *__restrict__ c, long long pitch, ZMap zm, int nx, int ny, int nz, const T *__restrict__ dinv,
               const TB *__restrict__ b, const T *__restrict__ wm, const T *__restrict__ wp,
               const T *__restrict__ yc, int cnx, int cny, int ncd, TO *__restrict__ y, T w, int kz,
               const int *skip) {
  if (skip && *skip) return;
  mgd::UpOp<T, D, TB, TO> op{c, pitch, nx, ny, dinv, b, wm, wp, yc, cnx, cny, ncd, y, w, zm};
  const int k0 = blockIdx.z * kz;
  march_tile<T, kTileX, TY>(op, nx, ny, nz, k0, min(k0 + kz, nz), blockIdx.x, blockIdx.y);
}

// Coarsest level: x = inv b (FP64 inverse, FP64 sums), one wavefront per row.
template <typename T>
__global__ void __launch_bounds__(256)
    k_coarse_gemv(const double *__restrict__ inv, int n, const T *__restrict__ b, T *__restrict__ x, const int *skip) {
  if (skip && *skip) return;
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= n) return;
  const double s = mgd::gemv_row<T>(inv, n, b, r, lane);
  if (lane == 0) x[r] = (T)s;
}

template <typename S, typename D>
__global__ void k_convert(const S *__restrict__ s, D *__restrict__ d, long long n, const int *skip) {
  if (skip && *skip) return;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
    d[t] = (D)s[t];
}

template <typename T>
__global__ void k_copy_t(const T *__restrict__ s, T *__restrict__ d, long long n, const int *skip) {
  if (skip && *skip) return;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
    d[t] = s[t];
}

unsigned nb(long long n) { return (unsigned)std::max(1LL, std::min((n + 255) / 256, 65536LL)); }

// ---------------------------------------------------------------- device hierarchy
struct DLevel {
  Dims d;
  int cdir = -1;
  double omega = 1.0;
  long long pitch = 0;
  void *coef = nullptr;  // owned unless level 0 borrows the matrix planes
  bool own_coef = true;
  void *dinv = nullptr;
  bool own_dinv = true;
  void *wm = nullptr, *wp = nullptr;
  void *x = nullptr, *y = nullptr, *b = nullptr, *r = nullptr, *tmp = nullptr;
  // What the fused down/up kernels read: hot copies of the coefficients, 1/diag
  // and weights when the level is z-extruded (zm.pl != 0; xf flag words for the
  // nexc flagged points), else aliases of the full arrays.  own_hc / own_hw: hc, hd / hwm, hwp
  // allocated for the level.
  void *hc = nullptr, *hd = nullptr, *hwm = nullptr, *hwp = nullptr;
  unsigned long long *xf = nullptr;
  long long hpitch = 0, nexc = 0;
  ZMap zm;
  bool own_hc = false, own_hw = false;
};

struct PfmgImpl {
  int maxit = 1;
  double tol = 0.0;
  int relax = 1, rap = 1, npre = 1, npost = 1, maxlev = 0, print = 0, bits = 64;
  long long coarse = 0, built_coarse = -1;  // coarsest-size cap; 0 = coarse_max(N0)
  bool zero_guess = false;
  int its = 0;
  fpc_StructMatrix_s *A = nullptr;
  unsigned long long version = ~0ULL;
  int built_bits = 0;
  std::vector<DLevel> L;
  double *inv = nullptr;
  int ncoarse = 0;
  double *w64 = nullptr;  // FP64 work (stand-alone solves)
  double *e64 = nullptr;
  // Captured V-cycles, one per (input, output, skip flag) seen ("mg_graph" tuning key).
  struct Graph {
    const double *in;
    double *out;
    const int *skip;
    int bits;
    hipGraphExec_t exec;
  };
  std::vector<Graph> graphs;
};

void free_graphs(PfmgImpl *P) {
  for (auto &g : P->graphs) (void)hipGraphExecDestroy(g.exec);
  P->graphs.clear();
}

void free_levels(PfmgImpl *P) {
  free_graphs(P);  // their launches hold the levels' buffers
  for (auto &l : P->L) {
    if (l.own_coef && l.coef) (void)hipFree(l.coef);
    if (l.own_dinv && l.dinv) (void)hipFree(l.dinv);
    for (void *q : {l.wm, l.wp, l.x, l.y, l.b, l.r, l.tmp})
      if (q) (void)hipFree(q);
    if (l.own_hc) {
      if (l.hc) (void)hipFree(l.hc);
      if (l.hd) (void)hipFree(l.hd);
    }
    if (l.own_hw) {
      if (l.hwm) (void)hipFree(l.hwm);
      if (l.hwp) (void)hipFree(l.hwp);
    }
    if (l.xf) (void)hipFree(l.xf);
  }
  P->L.clear();
  if (P->inv) (void)hipFree(P->inv);
  P->inv = nullptr;
  if (P->w64) (void)hipFree(P->w64);
  if (P->e64) (void)hipFree(P->e64);
  P->w64 = P->e64 = nullptr;
}

template <typename T>
int upload(void **dst, const std::vector<double> &src, size_t count, size_t n_alloc) {
  std::vector<T> tmp(n_alloc, (T)0);
  for (size_t q = 0; q < count; ++q) tmp[q] = (T)src[q];
  FPC_HIP(hipMalloc(dst, n_alloc * sizeof(T)));
  FPC_HIP(hipMemcpyAsync(*dst, tmp.data(), n_alloc * sizeof(T), hipMemcpyHostToDevice, stream()));
  FPC_HIP(hipStreamSynchronize(stream()));  // tmp is released on return
  return 0;
}

template <typename T>
int alloc_t(void **dst, long long n) {
  FPC_HIP(hipMalloc(dst, (size_t)std::max(1LL, n) * sizeof(T)));
  FPC_HIP(hipMemsetAsync(*dst, 0, (size_t)std::max(1LL, n) * sizeof(T), stream()));
  return 0;
}

template <typename T>
__global__ void k_dinv_t(const T *__restrict__ c, T *__restrict__ d, long long n) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
    d[t] = (c[t] != (T)0) ? (T)1 / c[t] : (T)0;
}

// The arrays the fused kernels of level l read (DLevel::hc .. zm): a hot copy
// (extrude.hpp) when the level's coefficients and weights repeat along z, else
// aliases of the full arrays.  1/diag is derived as the full one is: on the
// device from the coefficients at level 0 (as the matrix's), on the host at
// the levels built here.
template <typename T>
int build_hot(const HLevel &h, size_t l, DLevel &dl) {
  ExtrudePlan plan;
  // The weights as the hot copy holds them.  Only F points read weights; when the
  // level coarsens z its C planes (odd k) take the next F plane's weights so that
  // the interior planes repeat (their own values are never read).
  std::vector<double> wm, wp;
  if (h.cdir >= 0) {
    wm = h.wm;
    wp = h.wp;
    if (h.cdir == 2) {
      const size_t pl = (size_t)h.n[0] * (size_t)h.n[1];
      for (int k = 1; k < h.n[2]; k += 2) {
        const int src = (k + 1 < h.n[2]) ? k + 1 : k - 1;
        std::memcpy(wm.data() + (size_t)k * pl, wm.data() + (size_t)src * pl, pl * sizeof(double));
        std::memcpy(wp.data() + (size_t)k * pl, wp.data() + (size_t)src * pl, pl * sizeof(double));
      }
    }
  }
  if (z_extrude_enabled() && h.n[2] >= 5) {
    std::vector<const double *> arrays;
    for (int q = 0; q < kSlots; ++q) arrays.push_back(h.a.data() + (size_t)q * (size_t)h.N);
    if (h.cdir >= 0) {
      arrays.push_back(wm.data());
      arrays.push_back(wp.data());
    }
    plan = plan_extrusion(h.n[0], h.n[1], h.n[2], arrays);
  }
  if (!plan.m) {
    dl.hc = dl.coef; dl.hd = dl.dinv; dl.hwm = dl.wm; dl.hwp = dl.wp;
    dl.hpitch = dl.pitch;
    return 0;
  }
  int rc;
  const long long rows = plan.xoff + h.N;
  dl.hpitch = pad256(rows, (int)sizeof(T));
  std::vector<double> hot = hot_planes(h.a.data(), h.N, kSlots, h.n[0], h.n[1], h.n[2], plan, dl.hpitch);
  dl.own_hc = true;
  if ((rc = upload<T>(&dl.hc, hot, hot.size(), hot.size()))) return rc;
  if (l == 0) {
    FPC_HIP(hipMalloc(&dl.hd, (size_t)dl.hpitch * sizeof(T)));
    k_dinv_t<T><<<(unsigned)std::min((dl.hpitch + 255) / 256, 65536LL), 256, 0, stream()>>>((const T *)dl.hc,
                                                                                           (T *)dl.hd, dl.hpitch);
    FPC_LAUNCH_CHECK();
  } else {
    std::vector<double> dv((size_t)dl.hpitch, 0.0);
    for (size_t q = 0; q < dv.size(); ++q) dv[q] = (hot[q] != 0.0) ? 1.0 / hot[q] : 0.0;
    if ((rc = upload<T>(&dl.hd, dv, dv.size(), dv.size()))) return rc;
  }
  if (h.cdir >= 0) {
    dl.own_hw = true;
    std::vector<double> hwm = hot_planes(wm.data(), h.N, 1, h.n[0], h.n[1], h.n[2], plan, rows);
    std::vector<double> hwp = hot_planes(wp.data(), h.N, 1, h.n[0], h.n[1], h.n[2], plan, rows);
    if ((rc = upload<T>(&dl.hwm, hwm, hwm.size(), hwm.size()))) return rc;
    if ((rc = upload<T>(&dl.hwp, hwp, hwp.size(), hwp.size()))) return rc;
  }
  if (plan.nexc) {
    const std::vector<unsigned long long> words = flag_words(plan, h.n[0], h.n[1], h.n[2]);
    FPC_HIP(hipMalloc(&dl.xf, words.size() * sizeof(unsigned long long)));
    FPC_HIP(hipMemcpyAsync(dl.xf, words.data(), words.size() * sizeof(unsigned long long), hipMemcpyHostToDevice,
                           stream()));
    FPC_HIP(hipStreamSynchronize(stream()));  // words is released on return
  }
  dl.nexc = plan.nexc;
  dl.zm.pl = (long long)h.n[0] * h.n[1];
  dl.zm.nz = h.n[2];
  dl.zm.m = plan.m;
  dl.zm.xoff = plan.xoff;
  dl.zm.xb = dl.xf;
  return 0;
}

template <typename T>
int build_device(PfmgImpl *P, std::vector<HLevel> &H) {
  int rc;
  for (size_t l = 0; l < H.size(); ++l) {
    HLevel &h = H[l];
    DLevel dl;
    dl.d.nx = h.n[0]; dl.d.ny = h.n[1]; dl.d.nz = h.n[2]; dl.d.N = h.N;
    dl.cdir = h.cdir;
    dl.omega = h.omega;
    dl.pitch = pad256(std::max(1LL, h.N), (int)sizeof(T));
    if (l == 0) {
      // Level 0 borrows the matrix planes (FP64 or the FP32 copy).
      if (sizeof(T) == 8) {
        dl.coef = P->A->coef;
        dl.dinv = P->A->dinv;
        dl.pitch = P->A->pitch;
      } else {
        rc = ensure_fp32(P->A);
        if (rc) return rc;
        dl.coef = P->A->coef32;
        dl.dinv = P->A->dinv32;
        dl.pitch = P->A->pitch;
      }
      dl.own_coef = dl.own_dinv = false;
    } else {
      std::vector<double> planes((size_t)kSlots * (size_t)dl.pitch, 0.0), dv((size_t)dl.pitch, 0.0);
      for (int s = 0; s < kSlots; ++s)
        for (long long r = 0; r < h.N; ++r) planes[(size_t)s * (size_t)dl.pitch + (size_t)r] = h.A(s, r);
      for (long long r = 0; r < h.N; ++r) {
        double a0 = h.A(0, r);
        dv[(size_t)r] = (a0 != 0.0) ? 1.0 / a0 : 0.0;
      }
      if ((rc = upload<T>(&dl.coef, planes, planes.size(), planes.size()))) return rc;
      if ((rc = upload<T>(&dl.dinv, dv, dv.size(), dv.size()))) return rc;
    }
    if (h.cdir >= 0) {
      if ((rc = upload<T>(&dl.wm, h.wm, (size_t)h.N, (size_t)h.N))) return rc;
      if ((rc = upload<T>(&dl.wp, h.wp, (size_t)h.N, (size_t)h.N))) return rc;
    }
    if ((rc = alloc_t<T>(&dl.x, h.N))) return rc;
    if ((rc = alloc_t<T>(&dl.y, h.N))) return rc;
    if ((rc = alloc_t<T>(&dl.r, h.N))) return rc;
    if ((rc = alloc_t<T>(&dl.tmp, h.N))) return rc;
    if (l > 0 || sizeof(T) != 8)
      if ((rc = alloc_t<T>(&dl.b, h.N))) return rc;
    rc = build_hot<T>(h, l, dl);
    P->L.push_back(dl);  // owned allocations are freed with the level, also on error
    if (rc) return rc;
  }
  return 0;
}

int build(PfmgImpl *P, fpc_StructMatrix_s *A) {
  free_levels(P);
  P->A = A;
  int rc = ensure_dinv(A);
  if (rc) return rc;
  std::vector<HLevel> H(1);
  HLevel &h0 = H[0];
  h0.n[0] = A->grid.n[0]; h0.n[1] = A->grid.n[1]; h0.n[2] = A->grid.n[2];
  h0.N = A->grid.N;
  h0.a.assign((size_t)kSlots * (size_t)h0.N, 0.0);
  for (int s = 0; s < kSlots; ++s)
    FPC_HIP(hipMemcpyAsync(h0.a.data() + (size_t)s * (size_t)h0.N, A->coef + (long long)s * A->pitch,
                           (size_t)h0.N * sizeof(double), hipMemcpyDeviceToHost, stream()));
  FPC_HIP(hipStreamSynchronize(stream()));
  const long long cmax = P->coarse > 0 ? P->coarse : coarse_max(h0.N);
  for (;;) {
    HLevel &h = H.back();
    choose_direction(h);
    // MaxLevels only stops coarsening once the level is small enough for the
    // dense coarsest solve.
    bool lev_cap = P->maxlev > 0 && (int)H.size() >= P->maxlev && h.N <= cmax;
    if (h.N <= cmax || h.cdir < 0 || lev_cap) {
      h.cdir = -1;
      break;
    }
    interp_weights(h);
    HLevel c = coarsen(h);
    H.push_back(std::move(c));
  }
  std::vector<double> inv = dense_inverse(H.back());
  P->ncoarse = (int)H.back().N;
  if ((rc = upload<double>((void **)&P->inv, inv, inv.size(), std::max<size_t>(1, inv.size())))) return rc;
  rc = (P->bits == 32) ? build_device<float>(P, H) : build_device<double>(P, H);
  if (rc) return rc;
  P->version = A->version;
  P->built_bits = P->bits;
  P->built_coarse = P->coarse;
  return 0;
}

// What a fused kernel of level F reads: the hot arrays on z-marching launches of
// an extruded level, else the full arrays (one-point launches are latency-bound,
// and the flag a hot row may need would add a dependent load).
template <typename T>
struct HotView {
  const T *c, *d, *wm, *wp;
  long long pitch;
  ZMap zm;
  double rows;  // coefficient rows read
};
template <typename T>
HotView<T> hot_view(const DLevel &F) {
  if (F.zm.pl && tile_worthwhile(F.d))
    return HotView<T>{(const T *)F.hc, (const T *)F.hd, (const T *)F.hwm, (const T *)F.hwp, F.hpitch, F.zm,
                      stored_rows(F.zm, F.d.N, F.nexc)};
  return HotView<T>{(const T *)F.coef, (const T *)F.dinv, (const T *)F.wm, (const T *)F.wp, F.pitch, ZMap(),
                    (double)F.d.N};
}

// Tile rows of the down sweeps: 64x4 tiles (256 threads), so that with the registers of
// the pipelined march (march_tile_pipelined) three workgroups share a CU.
constexpr int kDownTY = 4;

// The fused down-sweep of level l: one weighted-Jacobi sweep from zero, the
// residual and its restriction into the coarse b.  b is read as TB (FP64 on the
// finest level of an FP32 cycle, rounded to T as it is read).
template <typename T, typename TB>
int fused_down(PfmgImpl *P, int l, const TB *b, T w, const int *skip, hipStream_t st) {
  DLevel &F = P->L[(size_t)l];
  DLevel &C = P->L[(size_t)l + 1];
  const double sz = (double)sizeof(T), szb = (double)sizeof(TB);
  // The z-marching kernels read the level's hot arrays (when it is extruded); the
  // one-point kernels of the latency-bound small levels the full ones.
  const HotView<T> H = hot_view<T>(F);
  const T *hc = H.c, *hd = H.d, *hwm = H.wm, *hwp = H.wp;
  const double rows = H.rows;  // rows of the coefficient, 1/diag and weight planes read
  // bytes: 7 coefficients and 1/diag as stored, b at every point, and r written -- or,
  // when z is restricted on the fly, one weight plane and the coarse b instead of r
  const char *cls = (l == 0) ? "mg_down0" : "mg_down";
  if (tile_worthwhile(F.d) && F.cdir == 2) {
    TileLaunch tl = tile_launch<kDownTY>(F.d);
    FPC_PROF_BEGIN(cls);
    k_down_restrict_z<T, kDownTY, TB><<<tl.grid, tl.block, 0, st>>>(hc, H.pitch, H.zm, F.d.nx, F.d.ny, F.d.nz, hd, b, hwm,
                                                               hwp, (T *)C.b, C.d.nz, w, tl.kz, skip);
    FPC_PROF_END(cls, rows * 9.0 * sz + (double)F.d.N * szb + (double)C.d.N * sz);
    FPC_LAUNCH_CHECK();
    return 0;
  }
  dim3 gc, bc;
  launch_shape(C.d, gc, bc, tile_tuning().mg_block);
  if (!tile_worthwhile(F.d)) {
    // small level: residual and restriction in one launch (the fine residual is not stored)
    FPC_PROF_BEGIN(cls);
    k_down_restrict<T, TB><<<gc, bc, 0, st>>>((const T *)F.coef, F.pitch, F.d.nx, F.d.ny, F.d.nz,
                                              (const T *)F.dinv, b, (const T *)F.wm, (const T *)F.wp, F.cdir, w,
                                              C.d.nx, C.d.ny, (T *)C.b, skip);
    FPC_PROF_END(cls, (double)F.d.N * (9.0 * sz + szb) + (double)C.d.N * sz);
    FPC_LAUNCH_CHECK();
    return 0;
  }
  FPC_PROF_BEGIN(cls);
  {
    TileLaunch tl = tile_launch<kDownTY>(F.d);
    k_down_tiled<T, kDownTY, TB><<<tl.grid, tl.block, 0, st>>>(hc, H.pitch, H.zm, F.d.nx, F.d.ny, F.d.nz, hd, b, (T *)F.r,
                                                         w, tl.kz, skip);
  }
  FPC_PROF_END(cls, rows * 8.0 * sz + (double)F.d.N * (szb + sz));
  FPC_LAUNCH_CHECK();
  {
    // bytes: r at every fine point, the weights at the F points, b of each coarse point
    FPC_PROF_BEGIN("mg_restrict");
    k_restrict<T><<<gc, bc, 0, st>>>(F.d.nx, F.d.ny, F.d.nz, F.cdir, (const T *)F.r, (const T *)F.wm,
                                    (const T *)F.wp, C.d.nx, C.d.ny, C.d.nz, (T *)C.b, skip);
    FPC_PROF_END("mg_restrict", ((double)F.d.N * 2.0 + (double)C.d.N) * sz);
  }
  FPC_LAUNCH_CHECK();
  return 0;
}

// The fused up-sweep of level l: prolongation of the coarse correction yc and one
// weighted-Jacobi sweep into y.  b is read as TB and y written as TO (FP64 on the
// finest level of an FP32 cycle).
template <typename T, typename TB, typename TO>
int fused_up(PfmgImpl *P, int l, const TB *b, const T *yc, TO *y, T w, const int *skip, hipStream_t st) {
  DLevel &F = P->L[(size_t)l];
  DLevel &C = P->L[(size_t)l + 1];
  // bytes: 7 coefficients, 1/diag and a weight pair as stored (the pair is read at
  // the F points, half of them), b at every point, y written; the coarse correction once
  const char *cls = (l == 0) ? "mg_up0" : "mg_up";
  const HotView<T> H = hot_view<T>(F);
  const double sz = (double)sizeof(T);
  const double bytes =
      H.rows * 9.0 * sz + (double)F.d.N * ((double)sizeof(TB) + (double)sizeof(TO)) + (double)C.d.N * sz;
  FPC_PROF_BEGIN(cls);
  const int ncd = (F.cdir == 0) ? C.d.nx : (F.cdir == 1 ? C.d.ny : C.d.nz);
  const T *hc = H.c, *hd = H.d;
  const T *wm = H.wm, *wp = H.wp;
  if (tile_worthwhile(F.d)) {
    TileLaunch tl = tile_launch<8>(F.d);
#define FPC_UP_TILED(D)                                                                                               \
  k_up_tiled<T, D, 8, TB, TO><<<tl.grid, tl.block, 0, st>>>(hc, H.pitch, H.zm, F.d.nx, F.d.ny, F.d.nz, hd, b, wm, wp, \
                                                            yc, C.d.nx, C.d.ny, ncd, y, w, tl.kz, skip)
    if (F.cdir == 0) FPC_UP_TILED(0);
    else if (F.cdir == 1) FPC_UP_TILED(1);
    else FPC_UP_TILED(2);
#undef FPC_UP_TILED
  } else {
    dim3 g, bl;
    launch_shape(F.d, g, bl, tile_tuning().mg_block);
#define FPC_UP_SIMPLE(D)                                                                                         \
  k_up_jacobi<T, D, TB, TO><<<g, bl, 0, st>>>((const T *)F.coef, F.pitch, F.d.nx, F.d.ny, F.d.nz,                  \
                                              (const T *)F.dinv, b, (const T *)F.wm, (const T *)F.wp, yc, C.d.nx, \
                                              C.d.ny, ncd, y, w, skip)
    if (F.cdir == 0) FPC_UP_SIMPLE(0);
    else if (F.cdir == 1) FPC_UP_SIMPLE(1);
    else FPC_UP_SIMPLE(2);
#undef FPC_UP_SIMPLE
  }
  FPC_PROF_END(cls, bytes);
  FPC_LAUNCH_CHECK();
  return 0;
}

// One V-cycle with zero initial guess: out = M^-1 in.
template <typename T>
int vcycle(PfmgImpl *P, const double *in, double *out, const int *skip) {
  hipStream_t st = stream();
  int nl = (int)P->L.size();
  bool fp64 = sizeof(T) == 8;
  DLevel &L0 = P->L[0];
  bool rb = P->relax >= 2;
  // one Jacobi sweep before and after: the fused down/up kernels
  bool fused = !rb && P->npre == 1 && P->npost == 1;
  // An FP32 cycle whose finest level runs the fused kernels reads `in` and writes
  // `out` (FP64) there directly, rounding as it reads and writes; otherwise the
  // FP64 vectors are converted to and from level-0 FP32 copies.
  const bool wide0 = !fp64 && fused && nl > 1 && tile_tuning().mg_wide_finest;
  const T *b0;
  if (fp64 || wide0) {
    b0 = (const T *)in;  // FP64 cycle: in itself; wide0: not read through this pointer
  } else {
    k_convert<double, T><<<nb(L0.d.N), 256, 0, st>>>(in, (T *)L0.b, L0.d.N, skip);
    FPC_LAUNCH_CHECK();
    b0 = (const T *)L0.b;
  }
  // level-0 solution buffer: the output itself in FP64
  T *y0 = fp64 ? (T *)out : (T *)L0.y;
  auto bof = [&](int l) -> const T * { return l == 0 ? b0 : (const T *)P->L[(size_t)l].b; };
  auto yof = [&](int l) -> T * { return l == 0 ? y0 : (T *)P->L[(size_t)l].y; };
  // red/black relaxes in place (x == y); Jacobi smooths x into y
  auto xof = [&](int l) -> T * { return rb ? yof(l) : (T *)P->L[(size_t)l].x; };
  auto w_of = [&](int l) -> T { return (T)(P->relax == 0 ? 1.0 : P->L[(size_t)l].omega); };
  for (int l = 0; l < nl - 1; ++l) {
    DLevel &F = P->L[(size_t)l];
    DLevel &C = P->L[(size_t)l + 1];
    dim3 g, bl;
    launch_shape(F.d, g, bl);
    const T *c = (const T *)F.coef;
    const T *dv = (const T *)F.dinv;
    T *x = xof(l);
    if (fused) {
      const int e = (l == 0 && wide0) ? fused_down<T, double>(P, 0, in, w_of(0), skip, st)
                                       : fused_down<T, T>(P, l, bof(l), w_of(l), skip, st);
      if (e) return e;
      continue;
    }
    if (rb) {
      for (int sw = 0; sw < P->npre; ++sw) {
        int c0 = 0, c1 = 1;
        k_relax_rb<T><<<g, bl, 0, st>>>(c, F.pitch, F.d.nx, F.d.ny, F.d.nz, dv, bof(l), x, c0, sw == 0, skip);
        k_relax_rb<T><<<g, bl, 0, st>>>(c, F.pitch, F.d.nx, F.d.ny, F.d.nz, dv, bof(l), x, c1, 0, skip);
      }
    } else {
      k_relax_zero<T><<<nb(F.d.N), 256, 0, st>>>(dv, bof(l), x, w_of(l), F.d.N, skip);
      for (int sw = 1; sw < P->npre; ++sw) {
        k_relax_jacobi<T><<<g, bl, 0, st>>>(c, F.pitch, F.d.nx, F.d.ny, F.d.nz, dv, bof(l), x, (T *)F.tmp, w_of(l), skip);
        k_copy_t<T><<<nb(F.d.N), 256, 0, st>>>((const T *)F.tmp, x, F.d.N, skip);
      }
    }
    FPC_LAUNCH_CHECK();
    k_residual_mg<T><<<g, bl, 0, st>>>(c, F.pitch, F.d.nx, F.d.ny, F.d.nz, bof(l), x, (T *)F.r, skip);
    dim3 gc, bc;
    launch_shape(C.d, gc, bc);
    k_restrict<T><<<gc, bc, 0, st>>>(F.d.nx, F.d.ny, F.d.nz, F.cdir, (const T *)F.r, (const T *)F.wm, (const T *)F.wp,
                                    C.d.nx, C.d.ny, C.d.nz, (T *)C.b, skip);
    FPC_LAUNCH_CHECK();
  }
  {
    int lc = nl - 1;
    k_coarse_gemv<T><<<(unsigned)((P->ncoarse + 3) / 4), 256, 0, st>>>(P->inv, P->ncoarse, bof(lc), yof(lc), skip);
    FPC_LAUNCH_CHECK();
  }
  for (int l = nl - 2; l >= 0; --l) {
    DLevel &F = P->L[(size_t)l];
    dim3 g, bl;
    launch_shape(F.d, g, bl);
    const T *c = (const T *)F.coef;
    const T *dv = (const T *)F.dinv;
    T *x = xof(l);
    if (fused) {
      const int e = (l == 0 && wide0)
                        ? fused_up<T, double, double>(P, 0, in, (const T *)yof(1), out, w_of(0), skip, st)
                        : fused_up<T, T, T>(P, l, bof(l), (const T *)yof(l + 1), yof(l), w_of(l), skip, st);
      if (e) return e;
      continue;
    }
    DLevel &C = P->L[(size_t)l + 1];
    k_prolong_add<T><<<g, bl, 0, st>>>(F.d.nx, F.d.ny, F.d.nz, F.cdir, (const T *)F.wm, (const T *)F.wp,
                                       (const T *)yof(l + 1), C.d.nx, C.d.ny, C.d.nz, x, skip);
    if (rb) {
      for (int sw = 0; sw < P->npost; ++sw) {
        int c0 = (P->relax == 2) ? 1 : 0, c1 = 1 - c0;
        k_relax_rb<T><<<g, bl, 0, st>>>(c, F.pitch, F.d.nx, F.d.ny, F.d.nz, dv, bof(l), x, c0, 0, skip);
        k_relax_rb<T><<<g, bl, 0, st>>>(c, F.pitch, F.d.nx, F.d.ny, F.d.nz, dv, bof(l), x, c1, 0, skip);
      }
      // x == y for red/black
    } else {
      T *y = yof(l);
      if (P->npost <= 0) {
        k_copy_t<T><<<nb(F.d.N), 256, 0, st>>>(x, y, F.d.N, skip);
      }
      for (int sw = 0; sw < P->npost; ++sw) {
        bool last = (sw == P->npost - 1);
        T *dst = last ? y : (T *)F.tmp;
        k_relax_jacobi<T><<<g, bl, 0, st>>>(c, F.pitch, F.d.nx, F.d.ny, F.d.nz, dv, bof(l), x, dst, w_of(l), skip);
        if (!last) k_copy_t<T><<<nb(F.d.N), 256, 0, st>>>((const T *)F.tmp, x, F.d.N, skip);
      }
    }
    FPC_LAUNCH_CHECK();
  }
  if (!fp64 && !wide0) {
    k_convert<T, double><<<nb(L0.d.N), 256, 0, st>>>((const T *)L0.y, out, L0.d.N, skip);
    FPC_LAUNCH_CHECK();
  }
  return 0;
}

// One V-cycle as a captured graph: the level launches of a V-cycle are fixed once the
// hierarchy is built, so the first cycle with a given input, output and skip flag is
// captured and later ones replay it (one launch call instead of one per kernel; the
// same kernels and arguments, so the same bits).  Not while kernel timing is on (its
// events would be captured), and at most 96 graphs per hierarchy.
template <typename T>
int vcycle_maybe_graph(PfmgImpl *P, const double *in, double *out, const int *skip) {
  if (!tile_tuning().mg_graph || prof_enabled()) return vcycle<T>(P, in, out, skip);
  const int bits = (int)sizeof(T) * 8;
  hipStream_t st = stream();
  for (const auto &g : P->graphs)
    if (g.in == in && g.out == out && g.skip == skip && g.bits == bits) {
      FPC_HIP(hipGraphLaunch(g.exec, st));
      return 0;
    }
  if (P->graphs.size() >= 96) return vcycle<T>(P, in, out, skip);
  FPC_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
  const int rc = vcycle<T>(P, in, out, skip);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &graph);
  if (rc || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    if (rc) return rc;
    FPC_HIP(e);
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  FPC_HIP(ei);
  P->graphs.push_back({in, out, skip, bits, exec});
  FPC_HIP(hipGraphLaunch(exec, st));
  return 0;
}

PfmgImpl *pimpl(fpc_StructSolver s) {
  if (!s || s->kind != SK_PFMG) return nullptr;
  return (PfmgImpl *)s->impl;
}

}  // namespace

int pfmg_setup(void *pf, fpc_StructMatrix_s *A) {
  PfmgImpl *P = (PfmgImpl *)pf;
  if (P->A == A && P->version == A->version && P->built_bits == P->bits && P->built_coarse == P->coarse &&
      !P->L.empty())
    return 0;
  return build(P, A);
}

int pfmg_batch_view(void *pf, std::vector<MgLevelView> &levels, const double **inv, int *ncoarse) {
  PfmgImpl *P = (PfmgImpl *)pf;
  if (!P || P->L.empty()) {
    set_error(FPC_ERROR_ARG, "batched solve: PFMG used before setup");
    return FPC_ERROR_ARG;
  }
  const bool fused = P->relax < 2 && P->npre == 1 && P->npost == 1;
  if (P->built_bits != 64 || !fused || P->maxit != 1) {
    set_error(FPC_ERROR_ARG,
              "batched solve: only the FP64 V-cycle with one (weighted) Jacobi sweep each way, one cycle, is supported");
    return FPC_ERROR_ARG;
  }
  levels.clear();
  for (const DLevel &L : P->L) {
    MgLevelView v;
    v.nx = L.d.nx;
    v.ny = L.d.ny;
    v.nz = L.d.nz;
    v.N = L.d.N;
    v.cdir = L.cdir;
    v.w = (P->relax == 0) ? 1.0 : L.omega;
    v.pitch = L.hpitch;
    v.zm = L.zm;
    v.coef = (const double *)L.hc;
    v.dinv = (const double *)L.hd;
    v.wm = (const double *)L.hwm;
    v.wp = (const double *)L.hwp;
    v.fpitch = L.pitch;
    v.fcoef = (const double *)L.coef;
    v.fdinv = (const double *)L.dinv;
    v.fwm = (const double *)L.wm;
    v.fwp = (const double *)L.wp;
    v.b = (double *)L.b;
    v.r = (double *)L.r;
    v.y = (double *)L.y;
    levels.push_back(v);
  }
  *inv = P->inv;
  *ncoarse = P->ncoarse;
  return 0;
}

int pfmg_apply(void *pf, const double *b, double *x, const int *skip) {
  PfmgImpl *P = (PfmgImpl *)pf;
  if (P->L.empty()) {
    set_error(FPC_ERROR_GENERIC, "PFMG used before setup");
    return FPC_ERROR_GENERIC;
  }
  int rc = (P->built_bits == 32) ? vcycle_maybe_graph<float>(P, b, x, skip) : vcycle_maybe_graph<double>(P, b, x, skip);
  if (rc) return rc;
  // further cycles as a stationary iteration (MaxIter > 1 as a preconditioner)
  for (int it = 1; it < P->maxit; ++it) {
    long long n = P->A->grid.N;
    if (!P->w64) FPC_HIP(hipMalloc(&P->w64, (size_t)std::max(1LL, n) * sizeof(double)));
    if (!P->e64) FPC_HIP(hipMalloc(&P->e64, (size_t)std::max(1LL, n) * sizeof(double)));
    if ((rc = residual_norm_partials(P->A, b, x, P->w64, scratch_doubles(kRedBlocks, 5)))) return rc;
    rc = (P->built_bits == 32) ? vcycle<float>(P, P->w64, P->e64, skip) : vcycle<double>(P, P->w64, P->e64, skip);
    if (rc) return rc;
    if ((rc = axpy_inplace(x, P->e64, n))) return rc;
  }
  return 0;
}

}  // namespace fpc

using namespace fpc;

extern "C" {

int fpc_StructPFMGCreate(fpc_StructSolver *solver) {
  FPC_CHECK_ARG(solver, "null handle pointer");
  ensure_init();
  fpc_StructSolver s = new fpc_StructSolver_s();
  s->kind = SK_PFMG;
  s->impl = new PfmgImpl();
  *solver = s;
  return 0;
}

#define PIMPL(s)                                  \
  PfmgImpl *P = pimpl(s);                         \
  FPC_CHECK_ARG(P, "not a PFMG solver handle")

int fpc_StructPFMGSetMaxIter(fpc_StructSolver s, int maxit) { PIMPL(s); P->maxit = std::max(1, maxit); return 0; }
int fpc_StructPFMGSetTol(fpc_StructSolver s, double tol) { PIMPL(s); P->tol = tol; return 0; }
int fpc_StructPFMGSetRelaxType(fpc_StructSolver s, int t) {
  PIMPL(s);
  FPC_CHECK_ARG(t >= 0 && t <= 3, "relax type must be 0..3");
  P->relax = t;
  return 0;
}
int fpc_StructPFMGSetRAPType(fpc_StructSolver s, int t) {
  PIMPL(s);
  FPC_CHECK_ARG(t == 0 || t == 1, "RAP type must be 0 or 1");
  P->rap = t;
  return 0;
}
int fpc_StructPFMGSetNumPreRelax(fpc_StructSolver s, int n) { PIMPL(s); P->npre = std::max(1, n); return 0; }
int fpc_StructPFMGSetNumPostRelax(fpc_StructSolver s, int n) { PIMPL(s); P->npost = std::max(0, n); return 0; }
int fpc_StructPFMGSetMaxLevels(fpc_StructSolver s, int n) { PIMPL(s); P->maxlev = std::max(0, n); return 0; }
int fpc_StructPFMGSetPrintLevel(fpc_StructSolver s, int level) { PIMPL(s); P->print = level; return 0; }
int fpc_StructPFMGSetZeroGuess(fpc_StructSolver s) { PIMPL(s); P->zero_guess = true; return 0; }
int fpc_StructPFMGSetNonZeroGuess(fpc_StructSolver s) { PIMPL(s); P->zero_guess = false; return 0; }
int fpc_StructPFMGSetPrecision(fpc_StructSolver s, int bits) {
  PIMPL(s);
  FPC_CHECK_ARG(bits == 32 || bits == 64, "precision must be 32 or 64");
  P->bits = bits;
  return 0;
}

int fpc_StructPFMGSetCoarseSize(fpc_StructSolver s, int n) {
  PIMPL(s);
  FPC_CHECK_ARG(n >= 0 && n <= kCoarseSizeLimit, "coarse size must be 0..4096");
  P->coarse = n;
  return 0;
}

int fpc_StructPFMGSetup(fpc_StructSolver s, fpc_StructMatrix A, fpc_StructVector b, fpc_StructVector x) {
  PIMPL(s);
  FPC_CHECK_ARG(A && A->initialized, "matrix not initialized");
  (void)b;
  (void)x;
  int rc = pfmg_setup(P, A);
  if (rc) return rc;
  FPC_HIP(hipStreamSynchronize(stream()));
  return 0;
}

// Stand-alone solve: MaxIter V-cycles (x <- x + M^-1 (b - A x)), stopping when
// ||r||/||b|| <= Tol if Tol > 0.
int fpc_StructPFMGSolve(fpc_StructSolver s, fpc_StructMatrix A, fpc_StructVector b, fpc_StructVector x) {
  PIMPL(s);
  FPC_CHECK_ARG(A && b && x && A->grid.N == b->grid.N && A->grid.N == x->grid.N, "bad arguments");
  int rc = pfmg_setup(P, A);
  if (rc) return rc;
  long long n = A->grid.N;
  if (!P->w64) FPC_HIP(hipMalloc(&P->w64, (size_t)std::max(1LL, n) * sizeof(double)));
  if (!P->e64) FPC_HIP(hipMalloc(&P->e64, (size_t)std::max(1LL, n) * sizeof(double)));
  double *part = scratch_doubles(kRedBlocks + 8, 5);
  if (!part) return FPC_ERROR_MEMORY;
  double bnorm2 = 0.0;
  if (P->tol > 0.0) {
    if ((rc = dot_partials(b->data, b->data, n, part)) || (rc = reduce_partials(part, red_blocks(n), part + kRedBlocks)))
      return rc;
    FPC_HIP(hipMemcpyAsync(&bnorm2, part + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, stream()));
    FPC_HIP(hipStreamSynchronize(stream()));
  }
  P->its = 0;
  for (int it = 0; it < P->maxit; ++it) {
    if (it == 0 && P->zero_guess) {
      rc = (P->built_bits == 32) ? vcycle<float>(P, b->data, x->data, nullptr) : vcycle<double>(P, b->data, x->data, nullptr);
      if (rc) return rc;
    } else {
      if ((rc = residual_norm_partials(A, b->data, x->data, P->w64, part))) return rc;
      rc = (P->built_bits == 32) ? vcycle<float>(P, P->w64, P->e64, nullptr) : vcycle<double>(P, P->w64, P->e64, nullptr);
      if (rc) return rc;
      if ((rc = axpy_inplace(x->data, P->e64, n))) return rc;
    }
    P->its = it + 1;
    if (P->tol > 0.0) {
      if ((rc = residual_norm_partials(A, b->data, x->data, nullptr, part)) ||
          (rc = reduce_partials(part, red_blocks(n), part + kRedBlocks)))
        return rc;
      double r2 = 0.0;
      FPC_HIP(hipMemcpyAsync(&r2, part + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, stream()));
      FPC_HIP(hipStreamSynchronize(stream()));
      if (std::sqrt(r2) <= P->tol * std::sqrt(bnorm2)) break;
    }
  }
  FPC_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int fpc_StructPFMGGetNumIterations(fpc_StructSolver s, int *n) {
  PIMPL(s);
  FPC_CHECK_ARG(n, "null pointer");
  *n = P->its;
  return 0;
}

int fpc_StructPFMGGetNumLevels(fpc_StructSolver s, int *n) {
  PIMPL(s);
  FPC_CHECK_ARG(n, "null pointer");
  *n = (int)P->L.size();
  return 0;
}

int fpc_StructPFMGGetLevelExtruded(fpc_StructSolver s, int level, int *wall_planes, long long *flagged) {
  PIMPL(s);
  FPC_CHECK_ARG(wall_planes && flagged && level >= 0 && level < (int)P->L.size(), "bad level");
  const DLevel &L = P->L[(size_t)level];
  *wall_planes = L.zm.pl ? L.zm.m : 0;
  *flagged = L.nexc;
  return 0;
}

int fpc_StructPFMGGetLevelInfo(fpc_StructSolver s, int level, int *dims, int *cdir, double *weight) {
  PIMPL(s);
  FPC_CHECK_ARG(level >= 0 && level < (int)P->L.size(), "level out of range");
  const DLevel &l = P->L[(size_t)level];
  if (dims) { dims[0] = l.d.nx; dims[1] = l.d.ny; dims[2] = l.d.nz; }
  if (cdir) *cdir = l.cdir;
  if (weight) *weight = l.omega;
  return 0;
}

int fpc_StructPFMGGetLevelMatrix(fpc_StructSolver s, int level, double *values) {
  PIMPL(s);
  FPC_CHECK_ARG(values, "null values");
  FPC_CHECK_ARG(level >= 0 && level < (int)P->L.size(), "level out of range");
  const DLevel &l = P->L[(size_t)level];
  long long N = l.d.N;
  std::vector<double> planes((size_t)kSlots * (size_t)N);
  for (int q = 0; q < kSlots; ++q) {
    if (P->built_bits == 32) {
      std::vector<float> t((size_t)N);
      FPC_HIP(hipMemcpyAsync(t.data(), (const float *)l.coef + (long long)q * l.pitch, (size_t)N * sizeof(float),
                             hipMemcpyDeviceToHost, stream()));
      FPC_HIP(hipStreamSynchronize(stream()));
      for (long long r = 0; r < N; ++r) planes[(size_t)q * (size_t)N + (size_t)r] = t[(size_t)r];
    } else {
      FPC_HIP(hipMemcpyAsync(planes.data() + (size_t)q * (size_t)N, (const double *)l.coef + (long long)q * l.pitch,
                             (size_t)N * sizeof(double), hipMemcpyDeviceToHost, stream()));
      FPC_HIP(hipStreamSynchronize(stream()));
    }
  }
  for (long long r = 0; r < N; ++r)
    for (int q = 0; q < kSlots; ++q) values[7 * r + q] = planes[(size_t)q * (size_t)N + (size_t)r];
  return 0;
}

int fpc_StructPFMGDestroy(fpc_StructSolver s) {
  if (!s) return 0;
  PfmgImpl *P = pimpl(s);
  if (P) {
    free_levels(P);
    delete P;
  }
  delete s;
  return 0;
}

}  // extern "C"

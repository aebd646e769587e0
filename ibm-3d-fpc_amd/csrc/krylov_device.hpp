// krylov_device.hpp -- device state and per-block bodies of the GMRES kernels
// (libfpc).
//
// One solver's kernels (krylov.hip, struct_mv.hip) and the batched kernels that
// advance many solvers in one launch (batched.hip) run these same bodies, so a
// system solved in a batch computes bit for bit what it computes alone.  Block
// index and block count are parameters: a batched launch maps blockIdx.y to
// the system and passes that system's own partition.
#pragma once

#include <hip/hip_runtime.h>

#include "fpc_internal.hpp"
#include "stencil_tile.hpp"

namespace fpc {
namespace kry {

constexpr int KM = kMaxKDim;
constexpr int NB = kRedBlocks;  // row stride of the partial-sum arrays

// Status record polled by the host (first, small part of the device state).
struct Status {
  int i;        // Arnoldi steps completed in the current cycle
  int done;     // 1: converged estimate, 2: cycle full or maxit reached, 3: solve finished
  int iter;     // total steps
  int maxit;
  int kdim;
  int pad;
  double r_norm;  // current residual estimate |rs[i]|
  double eps;
};

struct State {
  Status st;
  double sigma[KM + 1];          // scale of stored basis vector j (v_j = sigma_j p_j)
  double rs[KM + 1];
  double c[KM], s[KM];
  double y[KM + 1];              // combination coefficients
  double hh[(KM + 1) * KM];      // hh(row, col) = hh[col * (KM + 1) + row]
};

__device__ inline double &H(State *S, int row, int col) { return S->hh[col * (KM + 1) + row]; }

// (A z)_r at point (i, j, k) = row r, with z scaled by 1/diag when dinv is given.
// DT: the stored 1/diag (double, or float for the FP32 Jacobi); the products are FP64.
// c and dinv are read at the rows zm gives (M: ZMap or FullMap).  Every load is
// issued at once: a neighbour outside the box is read at the point itself and its
// term skipped by a select (skipped, not added as a*0: the tile forms' bits).
template <typename DT = double, class M>
__device__ inline double arnoldi_row(const double *__restrict__ c, long long pitch, M zm, int nx, int ny, int nz,
                                     const double *__restrict__ z, const DT *__restrict__ dinv, int i, int j, int k,
                                     long long r) {
  const long long pl = (long long)nx * ny;
  const bool xm = i > 0, xp = i < nx - 1, ym = j > 0, yp = j < ny - 1, zl = k > 0, zh = k < nz - 1;
  const long long q = zm(r, k);
  const long long r1 = xm ? r - 1 : r, r2 = xp ? r + 1 : r, r3 = ym ? r - nx : r, r4 = yp ? r + nx : r;
  const long long r5 = zl ? r - pl : r, r6 = zh ? r + pl : r;
  double v0 = z[r], v1 = z[r1], v2 = z[r2], v3 = z[r3], v4 = z[r4], v5 = z[r5], v6 = z[r6];
  if (dinv) {
    v0 = v0 * (double)dinv[q];
    v1 = v1 * (double)dinv[zm(r1, k)];
    v2 = v2 * (double)dinv[zm(r2, k)];
    v3 = v3 * (double)dinv[zm(r3, k)];
    v4 = v4 * (double)dinv[zm(r4, k)];
    v5 = v5 * (double)dinv[zm(r5, zl ? k - 1 : k)];
    v6 = v6 * (double)dinv[zm(r6, zh ? k + 1 : k)];
  }
  double s = c[q] * v0;
  s = xm ? s + c[pitch + q] * v1 : s;
  s = xp ? s + c[2 * pitch + q] * v2 : s;
  s = ym ? s + c[3 * pitch + q] * v3 : s;
  s = yp ? s + c[4 * pitch + q] * v4 : s;
  s = zl ? s + c[5 * pitch + q] * v5 : s;
  s = zh ? s + c[6 * pitch + q] * v6 : s;
  return s;
}

// Partial dots <p_j, w> for j = 0..cur, eight vectors per pass over w; block bid of nblk.
__device__ inline void multidot_body(const double *__restrict__ V, long long vp, long long n, State *S,
                                     double *__restrict__ part, int bid, int nblk) {
  __shared__ double lds[8 * 4];
  int cur = S->st.i;
  const double *w = V + (long long)(cur + 1) * vp;
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int j0 = 0; j0 <= cur; j0 += 8) {
    int nj = min(8, cur + 1 - j0);
    double acc[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] = 0.0;
    // The walk runs from the vectors' ends to their starts, two points per trip (twice the
    // loads in flight per thread): the orthogonalisation pass that follows walks from the
    // starts, where these reads leave the vectors in the last-level cache.
    const long long stride = (long long)nblk * blockDim.x;
    const long long t0 = bid * (long long)blockDim.x + threadIdx.x;
    long long m = (t0 < n) ? (n - 1 - t0) / stride : -1;  // index of this thread's last point
    for (; m >= 1; m -= 2) {
      const long long ta = t0 + m * stride, tb = ta - stride;
      const double wa = w[ta], wb = w[tb];
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (q < nj) {
          const double *vq = V + (long long)(j0 + q) * vp;
          acc[q] += vq[ta] * wa;
          acc[q] += vq[tb] * wb;
        }
    }
    if (m == 0) {
      const double wa = w[t0];
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (q < nj) acc[q] += V[(long long)(j0 + q) * vp + t0] * wa;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      double v = acc[q];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      acc[q] = v;
    }
    __syncthreads();
    if (lane == 0)
#pragma unroll
      for (int q = 0; q < 8; ++q) lds[q * 4 + wid] = acc[q];
    __syncthreads();
    if (threadIdx.x < nj) {
      int q = threadIdx.x;
      part[(long long)(j0 + q) * NB + bid] = ((lds[q * 4] + lds[q * 4 + 1]) + lds[q * 4 + 2]) + lds[q * 4 + 3];
    }
  }
}

// h_{jv,cur} = sigma_jv * sum of the nred partials of row jv (one block of 256).
// Returns at once, for the whole block, once the solve is done or j > the current index.
__device__ inline void reduce_h_body(State *S, const double *__restrict__ part, int nred, int jv) {
  __shared__ double lds[4];
  // the partials' addresses depend on neither the status nor the basis index:
  // the status (vector loads) and all NB / 256 partials of a thread are loaded
  // together (an index past nred reads entry 0 of the row, and its term is
  // skipped), summed in order, and only then is the status tested
  const int done = __hip_atomic_load(&S->st.done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int cur = __hip_atomic_load(&S->st.i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const double *row = part + (long long)jv * NB;
  double v[NB / 256];
#pragma unroll
  for (int r = 0; r < NB / 256; ++r) {
    const int t = (int)threadIdx.x + 256 * r;
    v[r] = row[t < nred ? t : 0];
  }
  double acc = 0.0;
#pragma unroll
  for (int r = 0; r < NB / 256; ++r) acc = ((int)threadIdx.x + 256 * r < nred) ? acc + v[r] : acc;
  pin(acc);
  if (done || jv > cur) return;
  double tot = block_sum<256>(acc, lds);
  if (threadIdx.x == 0) H(S, jv, cur) = S->sigma[jv] * tot;
}

// w -= sum_j h_j sigma_j p_j with partial sums of w^2; block bid of nblk.
__device__ inline void orth_update_body(double *__restrict__ V, long long vp, long long n, State *S,
                                        double *__restrict__ part, int bid, int nblk) {
  __shared__ double coef[KM + 1];
  __shared__ double lds[4];
  int cur = S->st.i;
  for (int q = threadIdx.x; q <= cur; q += blockDim.x) coef[q] = H(S, q, cur) * S->sigma[q];
  __syncthreads();
  double *w = V + (long long)(cur + 1) * vp;
  double acc = 0.0;
  const long long stride = (long long)nblk * blockDim.x;
  long long t = bid * (long long)blockDim.x + threadIdx.x;
  for (; t + stride < n; t += 2 * stride) {
    double va = w[t], vb = w[t + stride];
    // eight basis vectors' loads in flight per trip (the sums stay in order q = 0..cur)
#pragma unroll 8
    for (int q = 0; q <= cur; ++q) {
      const double *vq = V + (long long)q * vp;
      va -= coef[q] * vq[t];
      vb -= coef[q] * vq[t + stride];
    }
    w[t] = va;
    w[t + stride] = vb;
    acc += va * va;
    acc += vb * vb;
  }
  if (t < n) {
    double v = w[t];
    for (int q = 0; q <= cur; ++q) v -= coef[q] * V[(long long)q * vp + t];
    w[t] = v;
    acc += v * v;
  }
  double tot = block_sum<256>(acc, lds);
  if (threadIdx.x == 0) part[bid] = tot;
}

// Norm of the new vector, Givens update, convergence flag (HYPRE's Arnoldi
// tail); one block of 256, nred partials.
// Returns at once, for the whole block, once the solve is done.
__device__ inline void arnoldi_tail_body(State *S, const double *__restrict__ part, int nred) {
  __shared__ double lds[4];
  // The Hessenberg column and the rotations are staged in LDS by all threads
  // while the norm is reduced, then rotated there by thread 0: the same
  // operations in the same order as on the State, without a chain of dependent
  // global loads.  The status is loaded with the partials and tested after.
  __shared__ double hc[KM + 2], cs[KM + 1], sn[KM + 1];
  const int done = __hip_atomic_load(&S->st.done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int cur = __hip_atomic_load(&S->st.i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // HYPRE's i - 1
  double v[NB / 256];
#pragma unroll
  for (int r = 0; r < NB / 256; ++r) {
    const int t = (int)threadIdx.x + 256 * r;
    v[r] = part[t < nred ? t : 0];
  }
  double acc = 0.0;
#pragma unroll
  for (int r = 0; r < NB / 256; ++r) acc = ((int)threadIdx.x + 256 * r < nred) ? acc + v[r] : acc;
  pin(acc);
  if (done) return;
  const int i = cur + 1;
  for (int q = threadIdx.x; q < i; q += blockDim.x) hc[q] = H(S, q, cur);
  for (int q = threadIdx.x; q < cur; q += blockDim.x) {
    cs[q] = S->c[q];
    sn[q] = S->s[q];
  }
  double tot = block_sum<256>(acc, lds);  // its barriers also publish the staging
  if (threadIdx.x == 0) {
    double t = sqrt(tot);
    hc[i] = t;
    S->sigma[i] = (t != 0.0) ? 1.0 / t : 1.0;
    for (int j = 1; j < i; ++j) {
      double tt = hc[j - 1];
      hc[j - 1] = sn[j - 1] * hc[j] + cs[j - 1] * tt;
      hc[j] = -sn[j - 1] * tt + cs[j - 1] * hc[j];
    }
    double g = sqrt(hc[i] * hc[i] + hc[cur] * hc[cur]);
    if (g == 0.0) g = 2.220446049250313e-16;
    const double c = hc[cur] / g, s = hc[i] / g;
    S->c[cur] = c;
    S->s[cur] = s;
    const double rsc = S->rs[cur];
    S->rs[i] = -hc[i] * rsc / g;
    S->rs[cur] = c * rsc;
    hc[cur] = s * hc[i] + c * hc[cur];
    S->st.r_norm = fabs(S->rs[i]);
    S->st.i = i;
    S->st.iter += 1;
    if (S->st.r_norm <= S->st.eps) S->st.done = 1;
    else if (i >= S->st.kdim || S->st.iter >= S->st.maxit) S->st.done = 2;
  }
  __syncthreads();
  for (int q = threadIdx.x; q <= i; q += blockDim.x) H(S, q, cur) = hc[q];
}

// Upper-triangular solve; y_j = sigma_j * rs_j (coefficients of the stored p_j). One thread.
__device__ inline void backsolve_body(State *S) {
  int i = S->st.i;
  double rs[KM + 1];
  for (int q = 0; q <= i; ++q) rs[q] = S->rs[q];
  rs[i - 1] = rs[i - 1] / H(S, i - 1, i - 1);
  for (int k = i - 2; k >= 0; --k) {
    double t = 0.0;
    for (int j = k + 1; j < i; ++j) t -= H(S, k, j) * rs[j];
    t += rs[k];
    rs[k] = t / H(S, k, k);
  }
  for (int q = 0; q < i; ++q) S->y[q] = rs[q] * S->sigma[q];
}

// [dinv *] sum_j y_j P_j at point t (y staged in LDS, i terms).
template <typename DT = double>
__device__ inline double combine_point(const double *__restrict__ P, long long vp, long long t, const double *y, int i,
                                       const DT *__restrict__ dinv) {
  double v = y[i - 1] * P[(long long)(i - 1) * vp + t];
  for (int q = i - 2; q >= 0; --q) v += y[q] * P[(long long)q * vp + t];
  if (dinv) v *= (double)dinv[t];
  return v;
}

// HYPRE's residual at a full restart: rs_j rotated back, p_0 = sum_j rs_j v_j. One thread.
__device__ inline void restart_coefs_body(State *S) {
  int i = S->st.i;
  for (int j = i; j > 0; --j) {
    S->rs[j - 1] = -S->s[j - 1] * S->rs[j];
    S->rs[j] = S->c[j - 1] * S->rs[j];
  }
  for (int j = 0; j <= i; ++j) S->y[j] = S->rs[j] * S->sigma[j];
}

__device__ inline double restart_point(const double *__restrict__ V, long long vp, long long t, const double *y, int i) {
  double v = y[0] * V[t];
  for (int q = 1; q <= i; ++q) v += y[q] * V[(long long)q * vp + t];
  return v;
}

// Start of a cycle: i = 0, rs_0 = r_norm, sigma_0 = 1/r_norm. One thread.
__device__ inline void cycle_start_body(State *S, double r_norm, double eps, int iter, int maxit, int kdim) {
  S->st.i = 0;
  S->st.done = 0;
  S->st.iter = iter;
  S->st.maxit = maxit;
  S->st.kdim = kdim;
  S->st.r_norm = r_norm;
  S->st.eps = eps;
  S->rs[0] = r_norm;
  S->sigma[0] = 1.0 / r_norm;
}

// The loads of (b - A x) at point t = (i, j, k): the vector values first (neighbours at
// clamped indices; an absent one reads point t and is zeroed in residual_eval), then the
// coefficients of row u.
struct ResidualLoads {
  double a[7], v[7], b;
  bool in[6];
};

__device__ inline void residual_loads_x(ResidualLoads &L, int nx, int ny, int nz, long long pl,
                                        const double *__restrict__ b, const double *__restrict__ x, long long t, int i,
                                        int j, int k) {
  L.in[0] = i > 0, L.in[1] = i < nx - 1, L.in[2] = j > 0, L.in[3] = j < ny - 1, L.in[4] = k > 0, L.in[5] = k < nz - 1;
  const long long off[6] = {-1, 1, -(long long)nx, (long long)nx, -pl, pl};
  L.v[0] = x[t];
#pragma unroll
  for (int q = 0; q < 6; ++q) L.v[q + 1] = x[t + opaque(L.in[q] ? off[q] : 0LL)];
  L.b = b[t];
}

__device__ inline void residual_loads_a(ResidualLoads &L, const double *__restrict__ c, long long pitch, long long u) {
#pragma unroll
  for (int q = 0; q < 7; ++q) L.a[q] = c[q * pitch + u];
}

__device__ inline double residual_eval(const ResidualLoads &L) {
  double s = L.a[0] * L.v[0];
#pragma unroll
  for (int q = 0; q < 6; ++q) s = fma_if(L.in[q], s, L.a[q + 1], L.v[q + 1]);
  return L.b - s;
}

// r = b - A x (r optional) with the block's partial sum of r^2 into part[bid]; block bid of nblk.
// Each thread takes the points t, t + S, t + 2S, ... (S = nblk * blockDim.x) and sums their r^2 in
// that order; two of them per trip so that both points' loads are in flight together.  The point
// coordinates walk by the stride's own (i, j, k) instead of being divided out of t.
template <class M>
__device__ inline void residual_body(const double *__restrict__ c, long long pitch, M zm, int nx, int ny, int nz,
                                     const double *__restrict__ b, const double *__restrict__ x, double *__restrict__ r,
                                     double *__restrict__ part, int bid, int nblk) {
  __shared__ double lds[8];
  const long long N = (long long)nx * ny * nz, pl = (long long)nx * ny;
  const long long S = (long long)nblk * blockDim.x;
  const int si = (int)(S % nx), sj = (int)((S / nx) % ny), sk = (int)(S / pl);
  long long t = bid * (long long)blockDim.x + threadIdx.x;
  int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / pl);
  double acc = 0.0;
  for (; t < N; t += 2 * S) {
    int i1 = i, j1 = j, k1 = k;
    ijk_advance(i1, j1, k1, si, sj, sk, nx, ny);
    const bool two = t + S < N;
    const long long t1 = two ? t + S : t;
    if (!two) i1 = i, j1 = j, k1 = k;
    const long long tq[2] = {t, t1};
    const int kq[2] = {k, k1};
    unsigned char fq[2];
    long long uq[2];
    ResidualLoads L0, L1;
    zm.template flags<2>(tq, kq, fq);
    residual_loads_x(L0, nx, ny, nz, pl, b, x, t, i, j, k);
    residual_loads_x(L1, nx, ny, nz, pl, b, x, t1, i1, j1, k1);
    zm.template rows<2>(tq, kq, fq, uq);
    residual_loads_a(L0, c, pitch, uq[0]);
    residual_loads_a(L1, c, pitch, uq[1]);
    __builtin_amdgcn_sched_barrier(0);  // every load of both points issued before the first wait
    const double r0 = residual_eval(L0), r1 = residual_eval(L1);
    if (r) r[t] = r0;
    acc += r0 * r0;
    if (two) {
      if (r) r[t1] = r1;
      acc += r1 * r1;
    }
    i = i1, j = j1, k = k1;
    ijk_advance(i, j, k, si, sj, sk, nx, ny);
  }
  double tot = block_sum<256>(acc, lds);
  if (threadIdx.x == 0) part[bid] = tot;
}

__device__ inline void dot_body(const double *__restrict__ a, const double *__restrict__ b, long long n,
                                double *__restrict__ part, int bid, int nblk) {
  __shared__ double lds[8];
  double acc = 0.0;
  for (long long t = bid * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)nblk * blockDim.x)
    acc += a[t] * b[t];
  double tot = block_sum<256>(acc, lds);
  if (threadIdx.x == 0) part[bid] = tot;
}

// r = b with per-block partial sums of b^2: the residual of a zero initial guess.  Each
// thread visits t, t + S, t + 2S, ... as residual_body does two points per trip, so the
// partials equal residual_body's with x = 0 bit for bit (b - A 0 = b exactly).
__device__ inline void copy_dot_body(const double *__restrict__ b, double *__restrict__ r, long long n,
                                     double *__restrict__ part, int bid, int nblk) {
  __shared__ double lds[8];
  double acc = 0.0;
  for (long long t = bid * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)nblk * blockDim.x) {
    const double v = b[t];
    r[t] = v;
    acc += v * v;
  }
  double tot = block_sum<256>(acc, lds);
  if (threadIdx.x == 0) part[bid] = tot;
}

// Sum of n partials into *out (one block of 1024).
__device__ inline void reduce_body(const double *__restrict__ part, int n, double *__restrict__ out) {
  __shared__ double lds[16];
  double acc = 0.0;
  for (int t = threadIdx.x; t < n; t += blockDim.x) acc += part[t];
  double tot = block_sum<1024>(acc, lds);
  if (threadIdx.x == 0) *out = tot;
}

// The same matvec on z-marching tiles (each row of A and z read once).
template <bool kDinv, typename DT = double>
struct ArnoldiOp {
  static constexpr int kAhead = 1;  // prefetch distance of the march (stencil_tile.hpp)
  const double *c;
  long long pitch;
  int nx, ny;
  const double *z;
  const DT *dinv;
  double *w;
  double sg;
  ZMap zm;  // rows of c and dinv
  struct Raw {
    double z, dinv;
    long long u;  // row of the point's coefficients and 1/diag
  };
  __device__ long long at(int i, int j, int k) const { return ((long long)k * ny + j) * nx + i; }
  __device__ Raw fetch(int i, int j, int k, int f) const {
    const long long t = at(i, j, k), u = zm.row(t, k, f);
    return Raw{z[t], kDinv ? (double)dinv[u] : 1.0, u};
  }
  __device__ double finish(const Raw &r) const { return kDinv ? r.z * r.dinv : r.z; }
  __device__ void out(const Coef7<double> &a, const Raw &, int i, int j, int k, const Nbr<double> &nb) {
    w[at(i, j, k)] = sg * a.apply(nb);
  }
};

}  // namespace kry
}  // namespace fpc

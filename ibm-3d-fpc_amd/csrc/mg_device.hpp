// mg_device.hpp -- point bodies of the one-point-per-thread multigrid kernels
// and the z-marching tile operations of the same stages (libfpc).
//
// Coefficients, 1/diag and the interpolation weights of a point are read at the
// row a ZMap gives for that point (full storage, or the hot copy of a z-extruded
// level); vectors are always full.
//
// The level kernels of one hierarchy (pfmg.hip) and the batched kernels that run
// a level of many hierarchies in one launch (batched.hip) evaluate these same
// functions, so a V-cycle in a batch is bit for bit the V-cycle alone, whichever
// of the two kernel forms a level runs.
#pragma once

#include <hip/hip_runtime.h>

#include "stencil_tile.hpp"

namespace fpc {
namespace mgd {

// The one-point bodies below issue every load of a point at once: a neighbour
// outside the box is read at the point itself (a valid address) and its term is
// then skipped by a select, so no load waits behind a bounds branch.  Skipping
// (rather than adding a*0) keeps the sums bit for bit those of the tile forms.

// s + a v when the neighbour is in the box, else s.
template <typename T>
__device__ inline T add_if(bool in, T s, T a, T v) {
  return in ? s + a * v : s;
}

// v unchanged, passed through a volatile no-op: memory operations cannot move
// across it, and what is computed from the result comes after it.  Applied to every
// raw loaded value of a one-point body after all of its loads, it makes the wave issue
// every load before the first wait (one memory round trip), instead of the scheduler
// interleaving arithmetic, and the waits it needs, between the loads.
template <typename T>
__device__ inline void hold(T &v) {
  asm volatile("" : "+v"(v));
}

// The loads of one Jacobi-from-zero residual (down_point): b (as stored, TB) and
// 1/diag at the point and its six neighbours (slot order C, -x, +x, -y, +y, -z, +z)
// and the seven coefficients of its row.  A neighbour outside the box is read at
// the point itself.
template <typename T, typename TB = T>
struct DownRaw {
  TB b[7];
  T d[7], a[7];
};

template <typename T, class M, typename TB = T>
__device__ inline DownRaw<T, TB> down_loads(const T *__restrict__ c, long long pitch, M zm, int nx, int ny, int nz,
                                            const T *__restrict__ dinv, const TB *__restrict__ b, int i, int j,
                                            int k, long long t) {
  const long long pl = (long long)nx * ny;
  const bool xm = i > 0, xp = i < nx - 1, ym = j > 0, yp = j < ny - 1, zl = k > 0, zh = k < nz - 1;
  const long long u = zm(t, k);
  const long long t1 = xm ? t - 1 : t, t2 = xp ? t + 1 : t, t3 = ym ? t - nx : t, t4 = yp ? t + nx : t;
  const long long t5 = zl ? t - pl : t, t6 = zh ? t + pl : t;
  DownRaw<T, TB> r;
  r.b[0] = b[t];
  r.d[0] = dinv[u];
  r.b[1] = b[t1];
  r.d[1] = dinv[zm(t1, k)];
  r.b[2] = b[t2];
  r.d[2] = dinv[zm(t2, k)];
  r.b[3] = b[t3];
  r.d[3] = dinv[zm(t3, k)];
  r.b[4] = b[t4];
  r.d[4] = dinv[zm(t4, k)];
  r.b[5] = b[t5];
  r.d[5] = dinv[zm(t5, zl ? k - 1 : k)];
  r.b[6] = b[t6];
  r.d[6] = dinv[zm(t6, zh ? k + 1 : k)];
#pragma unroll
  for (int s = 0; s < 7; ++s) r.a[s] = c[s * pitch + u];
  return r;
}

template <typename T, typename TB>
__device__ inline void hold_raw(DownRaw<T, TB> &r) {
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    hold(r.b[q]);
    hold(r.d[q]);
    hold(r.a[q]);
  }
}

// b - A (w D^-1 b) at (i, j, k) from its loads.
// b is rounded to T here (the value down_point has always used).
template <typename T, typename TB>
__device__ inline T down_eval(const DownRaw<T, TB> &r, T w, int i, int j, int k, int nx, int ny, int nz) {
  const bool xm = i > 0, xp = i < nx - 1, ym = j > 0, yp = j < ny - 1, zl = k > 0, zh = k < nz - 1;
  const T b0 = (T)r.b[0];
  const T x0 = w * b0 * r.d[0];
  const T x1 = w * (T)r.b[1] * r.d[1], x2 = w * (T)r.b[2] * r.d[2];
  const T x3 = w * (T)r.b[3] * r.d[3], x4 = w * (T)r.b[4] * r.d[4];
  const T x5 = w * (T)r.b[5] * r.d[5], x6 = w * (T)r.b[6] * r.d[6];
  T s = r.a[0] * x0;
  s = add_if(xm, s, r.a[1], x1);
  s = add_if(xp, s, r.a[2], x2);
  s = add_if(ym, s, r.a[3], x3);
  s = add_if(yp, s, r.a[4], x4);
  s = add_if(zl, s, r.a[5], x5);
  s = add_if(zh, s, r.a[6], x6);
  return b0 - s;
}

// Residual of one weighted-Jacobi sweep from zero: b - A (w D^-1 b) at (i, j, k) = row t.
template <typename T, class M, typename TB = T>
__device__ inline T down_point(const T *__restrict__ c, long long pitch, M zm, int nx, int ny, int nz,
                               const T *__restrict__ dinv, const TB *__restrict__ b, T w, int i, int j, int k,
                               long long t) {
  return down_eval<T, TB>(down_loads<T, M, TB>(c, pitch, zm, nx, ny, nz, dinv, b, i, j, k, t), w, i, j, k, nx, ny,
                          nz);
}

// Restriction to coarse point (i, j, k) of a level coarsened in direction d:
// r(C) + w+(F-) r(F-) + w-(F+) r(F+), C = fine 2c+1.
template <typename T, class M>
__device__ inline T restrict_point(int fnx, int fny, int fnz, int d, const T *__restrict__ r, const T *__restrict__ wm,
                                   const T *__restrict__ wp, M zm, int i, int j, int k) {
  const int fi = (d == 0) ? 2 * i + 1 : i, fj = (d == 1) ? 2 * j + 1 : j, fk = (d == 2) ? 2 * k + 1 : k;
  const int pd = (d == 0) ? fi : (d == 1 ? fj : fk);
  const int fnd = (d == 0) ? fnx : (d == 1 ? fny : fnz);
  const long long stride = (d == 0) ? 1 : (d == 1 ? (long long)fnx : (long long)fnx * fny);
  const long long f0 = ((long long)fk * fny + fj) * fnx + fi;
  const int dk = (d == 2) ? 1 : 0;
  const bool hp = pd + 1 < fnd;
  const long long fp = hp ? f0 + stride : f0;
  const T v = r[f0] + wp[zm(f0 - stride, fk - dk)] * r[f0 - stride];
  return add_if(hp, v, wm[zm(fp, hp ? fk + dk : fk)], r[fp]);
}

// The same restriction with the fine residuals (down_point) evaluated on the fly
// instead of read back: one launch less per level for the latency-bound small
// levels.  Fine box nx x ny x nz; coarse point (i, j, k).  Sums as restrict_point.
template <typename T, class M, typename TB = T>
__device__ inline T down_restrict_point(const T *__restrict__ c, long long pitch, M zm, int nx, int ny, int nz,
                                        const T *__restrict__ dinv, const TB *__restrict__ b, T w,
                                        const T *__restrict__ wm, const T *__restrict__ wp, int d, int i, int j,
                                        int k) {
  const int di = (d == 0), dj = (d == 1), dk = (d == 2);
  const int fi = di ? 2 * i + 1 : i, fj = dj ? 2 * j + 1 : j, fk = dk ? 2 * k + 1 : k;
  const int pd = di ? fi : (dj ? fj : fk);
  const int fnd = di ? nx : (dj ? ny : nz);
  const long long stride = di ? 1 : (dj ? (long long)nx : (long long)nx * ny);
  const long long f0 = ((long long)fk * ny + fj) * nx + fi;
  const bool hp = pd + 1 < fnd;
  const int pi = hp ? fi + di : fi, pj = hp ? fj + dj : fj, pk = hp ? fk + dk : fk;
  const long long fp = hp ? f0 + stride : f0;
  // every load of the three fine points and the two weights issued before the first wait
  DownRaw<T, TB> q0 = down_loads<T, M, TB>(c, pitch, zm, nx, ny, nz, dinv, b, fi, fj, fk, f0);
  DownRaw<T, TB> qm = down_loads<T, M, TB>(c, pitch, zm, nx, ny, nz, dinv, b, fi - di, fj - dj, fk - dk, f0 - stride);
  DownRaw<T, TB> qp = down_loads<T, M, TB>(c, pitch, zm, nx, ny, nz, dinv, b, pi, pj, pk, fp);
  T wpm = wp[zm(f0 - stride, fk - dk)], wmp = wm[zm(fp, pk)];
  hold_raw(q0);
  hold_raw(qm);
  hold_raw(qp);
  hold(wpm);
  hold(wmp);
  const T r0 = down_eval<T, TB>(q0, w, fi, fj, fk, nx, ny, nz);
  const T rm = down_eval<T, TB>(qm, w, fi - di, fj - dj, fk - dk, nx, ny, nz);
  const T rp = down_eval<T, TB>(qp, w, pi, pj, pk, nx, ny, nz);
  const T v = r0 + wpm * rm;
  return add_if(hp, v, wmp, rp);
}

// What the interpolated value x of fine point (i, j, k) needs, the level coarsened
// in direction D: C points (odd) inject e1; F points take wm e0 + wp e1, an
// absent coarse neighbour reading as a zero weight and value.
template <typename T>
struct UpLoads {
  T b, dinv, wm, wp, e0, e1;
  bool f;
  long long u;  // row of the point's coefficients
};
// kSpec: load every operand unconditionally at valid addresses and select (one
// load round trip, for the latency-bound one-point kernels), or load only what
// the point uses (the tile forms, which prefetch ahead); the values are the same.
template <typename T, int D, bool kSpec, class M, typename TB = T>
__device__ inline UpLoads<T> up_loads(const TB *b, const T *dinv, const T *wm, const T *wp, const T *yc, M zm, int i,
                                      int j, int k, int fnx, int fny, int cnx, int cny, int ncd) {
  const long long t = ((long long)k * fny + j) * fnx + i, u = zm(t, k);
  const int pd = (D == 0) ? i : (D == 1 ? j : k);
  const long long cstride = (D == 0) ? 1 : (D == 1 ? (long long)cnx : (long long)cnx * cny);
  const int q = (pd & 1) ? ((pd - 1) >> 1) : (pd >> 1);
  const int ci = (D == 0) ? q : i, cj = (D == 1) ? q : j, ck = (D == 2) ? q : k;
  const long long cr = ((long long)ck * cny + cj) * cnx + ci;
  const bool odd = pd & 1, lo = !odd && q >= 1, hi = !odd && q < ncd;
  UpLoads<T> r;
  r.u = u;
  r.b = (T)b[t];
  r.dinv = dinv[u];
  r.f = !odd;
  if (kSpec) {
    // valid coarse rows: cr exists for C points and for q < ncd; q = ncd implies q >= 1
    const long long c1 = (odd || hi) ? cr : cr - cstride;
    const long long c0 = lo ? cr - cstride : c1;
    const T wml = wm[u], wpl = wp[u], y0 = yc[c0], y1 = yc[c1];
    r.wm = lo ? wml : T(0);
    r.e0 = lo ? y0 : T(0);
    r.wp = hi ? wpl : T(0);
    r.e1 = (odd || hi) ? y1 : T(0);
  } else if (odd) {
    r.wm = r.wp = r.e0 = T(0);
    r.e1 = yc[cr];
  } else {
    r.wm = lo ? wm[u] : T(0);
    r.e0 = lo ? yc[cr - cstride] : T(0);
    r.wp = hi ? wp[u] : T(0);
    r.e1 = hi ? yc[cr] : T(0);
  }
  return r;
}
// The one-point form of up_loads (kSpec) in two parts: every operand loaded
// unconditionally at a valid address (up_raw), then the selects (up_select), so that
// the raw values can be held (hold) between the two: the same values as up_loads<kSpec>.
template <typename T, typename TB = T>
struct UpRaw {
  TB b;
  T dinv, wm, wp, y0, y1;
  bool odd, lo, hi;
  long long u;
};
template <typename T, int D, class M, typename TB = T>
__device__ inline UpRaw<T, TB> up_raw(const TB *b, const T *dinv, const T *wm, const T *wp, const T *yc, M zm, int i,
                                      int j, int k, int fnx, int fny, int cnx, int cny, int ncd) {
  const long long t = ((long long)k * fny + j) * fnx + i, u = zm(t, k);
  const int pd = (D == 0) ? i : (D == 1 ? j : k);
  const long long cstride = (D == 0) ? 1 : (D == 1 ? (long long)cnx : (long long)cnx * cny);
  const int q = (pd & 1) ? ((pd - 1) >> 1) : (pd >> 1);
  const int ci = (D == 0) ? q : i, cj = (D == 1) ? q : j, ck = (D == 2) ? q : k;
  const long long cr = ((long long)ck * cny + cj) * cnx + ci;
  UpRaw<T, TB> r;
  r.odd = pd & 1;
  r.lo = !r.odd && q >= 1;
  r.hi = !r.odd && q < ncd;
  r.u = u;
  // valid coarse rows: cr exists for C points and for q < ncd; q = ncd implies q >= 1
  const long long c1 = (r.odd || r.hi) ? cr : cr - cstride;
  const long long c0 = r.lo ? cr - cstride : c1;
  r.b = b[t];
  r.dinv = dinv[u];
  r.wm = wm[u];
  r.wp = wp[u];
  r.y0 = yc[c0];
  r.y1 = yc[c1];
  return r;
}
template <typename T, typename TB>
__device__ inline void hold_raw(UpRaw<T, TB> &r) {
  hold(r.b);
  hold(r.dinv);
  hold(r.wm);
  hold(r.wp);
  hold(r.y0);
  hold(r.y1);
}
template <typename T, typename TB>
__device__ inline UpLoads<T> up_select(const UpRaw<T, TB> &r) {
  UpLoads<T> o;
  o.u = r.u;
  o.b = (T)r.b;
  o.dinv = r.dinv;
  o.f = !r.odd;
  o.wm = r.lo ? r.wm : T(0);
  o.e0 = r.lo ? r.y0 : T(0);
  o.wp = r.hi ? r.wp : T(0);
  o.e1 = (r.odd || r.hi) ? r.y1 : T(0);
  return o;
}

// x = w D^-1 b plus the interpolated coarse correction: (0 + wm e0) + wp e1 at F
// points, e1 at C points.  The multiply-adds are written out as fused operations, so
// every kernel that evaluates this rounds it the same way, whatever contractions the
// code around it would let the compiler choose.
template <typename T>
__device__ inline T up_value(const UpLoads<T> &r, T w) {
  const T s = fma(r.wp, r.e1, fma(r.wm, r.e0, T(0)));
  return fma(w * r.b, r.dinv, r.f ? s : r.e1);
}

// The Jacobi sweep's output at a point: x + w (b - (A x)) / diag, as one fused step.
template <typename T>
__device__ inline T up_out(T x, T b, T ax, T dinv, T w) {
  return fma(w * (b - ax), dinv, x);
}

// Prolongation + one weighted-Jacobi sweep at (i, j, k) = row t:
// y = x + w D^-1 (b - A x) with x = up_value() at the point and its neighbours.
template <typename T, int D, class M, typename TB = T>
__device__ inline T up_point(const T *__restrict__ c, long long pitch, M zm, int nx, int ny, int nz,
                             const T *__restrict__ dinv, const TB *__restrict__ b, const T *__restrict__ wm,
                             const T *__restrict__ wp, const T *__restrict__ yc, int cnx, int cny, int ncd, T w, int i,
                             int j, int k, long long t) {
#define FPC_L(ii, jj, kk) up_raw<T, D, M, TB>(b, dinv, wm, wp, yc, zm, ii, jj, kk, nx, ny, cnx, cny, ncd)
  const bool xm = i > 0, xp = i < nx - 1, ym = j > 0, yp = j < ny - 1, zl = k > 0, zh = k < nz - 1;
  const long long u = zm(t, k);
  // every load of the point, its six neighbours and its row issued before the first wait
  UpRaw<T, TB> r0 = FPC_L(i, j, k);
  UpRaw<T, TB> r1 = FPC_L(xm ? i - 1 : i, j, k), r2 = FPC_L(xp ? i + 1 : i, j, k);
  UpRaw<T, TB> r3 = FPC_L(i, ym ? j - 1 : j, k), r4 = FPC_L(i, yp ? j + 1 : j, k);
  UpRaw<T, TB> r5 = FPC_L(i, j, zl ? k - 1 : k), r6 = FPC_L(i, j, zh ? k + 1 : k);
#undef FPC_L
  Coef7<T> a;
  a.load(c, pitch, u);
  hold_raw(r0);
  hold_raw(r1);
  hold_raw(r2);
  hold_raw(r3);
  hold_raw(r4);
  hold_raw(r5);
  hold_raw(r6);
#pragma unroll
  for (int q = 0; q < 7; ++q) hold(a.a[q]);
  const UpLoads<T> lc = up_select(r0), l1 = up_select(r1), l2 = up_select(r2), l3 = up_select(r3);
  const UpLoads<T> l4 = up_select(r4), l5 = up_select(r5), l6 = up_select(r6);
  const T xc = up_value<T>(lc, w);
  const T x1 = up_value<T>(l1, w), x2 = up_value<T>(l2, w);
  const T x3 = up_value<T>(l3, w), x4 = up_value<T>(l4, w);
  const T x5 = up_value<T>(l5, w), x6 = up_value<T>(l6, w);
  T s = a.a[0] * xc;
  s = add_if(xm, s, a.a[1], x1);
  s = add_if(xp, s, a.a[2], x2);
  s = add_if(ym, s, a.a[3], x3);
  s = add_if(yp, s, a.a[4], x4);
  s = add_if(zl, s, a.a[5], x5);
  s = add_if(zh, s, a.a[6], x6);
  return up_out<T>(xc, lc.b, s, lc.dinv, w);
}

// Row r of the coarsest solve x = inv b (FP64 sums), one wavefront; valid in lane 0.
// Four independent partial sums per lane keep four row loads in flight.
template <typename T>
__device__ inline double gemv_row(const double *__restrict__ inv, int n, const T *__restrict__ b, int r, int lane) {
  const double *row = inv + (size_t)r * n;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int q = lane;
  for (; q + 192 < n; q += 256) {
    s0 += row[q] * (double)b[q];
    s1 += row[q + 64] * (double)b[q + 64];
    s2 += row[q + 128] * (double)b[q + 128];
    s3 += row[q + 192] * (double)b[q + 192];
  }
  for (; q < n; q += 64) s0 += row[q] * (double)b[q];
  double s = (s0 + s1) + (s2 + s3);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  return s;
}

// ---------------------------------------------------------------- z-marching forms
// A map that answers with a row already known (the tile's fetch knows the flag).
struct RowOf {
  long long u;
  __device__ inline long long operator()(long long, int) const { return u; }
};

// The same arithmetic on the z-marching tiles (stencil_tile.hpp): every row read
// once, z-neighbours carried in registers; bit for bit the one-point results.

// Jacobi sweep from zero and residual: r = b - A (w D^-1 b).  b is stored as TB
// (wider than T on the finest level of a reduced-precision cycle).
template <typename T, typename TB = T>
struct DownOp {
  static constexpr int kAhead = 2;           // prefetch distance of march_tile (stencil_tile.hpp)
  static constexpr int kAheadPipelined = 1;  // and of march_tile_pipelined
  const T *c;
  long long pitch;
  int nx, ny;
  const T *dinv;
  const TB *b;
  T *r;
  T w;
  ZMap zm;  // rows of c and dinv
  struct Raw {
    T b, dinv;
    long long u;
  };
  __device__ long long at(int i, int j, int k) const { return ((long long)k * ny + j) * nx + i; }
  __device__ Raw fetch(int i, int j, int k, int f) const {
    const long long t = at(i, j, k), u = zm.row(t, k, f);
    return Raw{(T)b[t], dinv[u], u};
  }
  __device__ T finish(const Raw &q) const { return w * q.b * q.dinv; }
  __device__ void out(const Coef7<T> &a, const Raw &own, int i, int j, int k, const Nbr<T> &nb) {
    r[at(i, j, k)] = own.b - a.apply(nb);
  }
};

// The same, restricted on the fly when the level coarsens z: coarse plane kc
// (fine C plane 2kc+1) = r(C) + wp(F-) r(F-) + wm(F+) r(F+), in that order; the
// fine residual is never stored.  Chunks start on even planes and run one plane
// past their end so the F+ plane of their last coarse plane is seen.  The weights
// of a point come with its fetch, ahead of use.
template <typename T, typename TB = T>
struct DownRestrictZOp {
  static constexpr int kAhead = 2;           // prefetch distance of march_tile (stencil_tile.hpp)
  static constexpr int kAheadPipelined = 1;  // and of march_tile_pipelined
  const T *c;
  long long pitch;
  int nx, ny;
  const T *dinv;
  const TB *b;
  T w;
  ZMap zm;  // rows of c, dinv, wm and wp
  const T *wm, *wp;
  T *bc;
  int nz, ncz, k0;
  T r_prev = T(0), wp_prev = T(0), pend = T(0);  // the plane below: residual, its wp
  struct Raw {
    T b, dinv, wm, wp;
    long long u;
  };
  __device__ long long at(int i, int j, int k) const { return ((long long)k * ny + j) * nx + i; }
  __device__ Raw fetch(int i, int j, int k, int f) const {
    const long long t = at(i, j, k), u = zm.row(t, k, f);
    return Raw{(T)b[t], dinv[u], wm[u], wp[u], u};
  }
  __device__ T finish(const Raw &q) const { return w * q.b * q.dinv; }
  __device__ void out(const Coef7<T> &a, const Raw &own, int i, int j, int k, const Nbr<T> &nb) {
    const T rk = own.b - a.apply(nb);
    if (k & 1) {
      const T v = rk + wp_prev * r_prev;
      const int kc = (k - 1) >> 1;
      if (k + 1 < nz) pend = v;
      else if (kc < ncz) bc[((long long)kc * ny + j) * nx + i] = v;
    } else if (k > k0) {
      const int kc = (k >> 1) - 1;
      if (kc < ncz) bc[((long long)kc * ny + j) * nx + i] = pend + own.wm * rk;
    }
    r_prev = rk;
    wp_prev = own.wp;
  }
};

// Prolongation of the coarse correction plus one Jacobi sweep:
// y = x + P e, then y + w D^-1 (b - A y), with x = w D^-1 b recomputed.  b is
// stored as TB and y written as TO (both wider than T on the finest level of a
// reduced-precision cycle: the rounding of b to T and of y from T happen here).
template <typename T, int D, typename TB = T, typename TO = T>
struct UpOp {
  static constexpr int kAhead = 1;  // prefetch distance of the march (stencil_tile.hpp)
  const T *c;
  long long pitch;
  int nx, ny;
  const T *dinv;
  const TB *b;
  const T *wm, *wp, *yc;
  int cnx, cny, ncd;
  TO *y;
  T w;
  ZMap zm;  // rows of c, dinv, wm and wp
  using Raw = UpLoads<T>;
  __device__ long long at(int i, int j, int k) const { return ((long long)k * ny + j) * nx + i; }
  __device__ Raw fetch(int i, int j, int k, int f) const {
    const RowOf fixed{zm.row(at(i, j, k), k, f)};
    return up_loads<T, D, false, RowOf, TB>(b, dinv, wm, wp, yc, fixed, i, j, k, nx, ny, cnx, cny, ncd);
  }
  __device__ T finish(const Raw &r) const { return up_value<T>(r, w); }
  __device__ void out(const Coef7<T> &a, const Raw &own, int i, int j, int k, const Nbr<T> &nb) {
    y[at(i, j, k)] = (TO)up_out<T>(nb.c, own.b, a.apply(nb), own.dinv, w);
  }
};

}  // namespace mgd
}  // namespace fpc

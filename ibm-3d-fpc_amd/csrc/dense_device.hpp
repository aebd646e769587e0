// dense_device.hpp -- device helpers of the staggered-field operations (libfpc):
// full-array index maps (x fastest, common.hpp:230-238 of the reference), the
// centred convection at a face (gpu_dense_ops.hpp:1-14), G p at a face in CSR
// column order and D q at a cell in CSR column order.  Shared by the forward
// stages (dense_ops.hip) and their adjoints (adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "fpc_dense.h"
#include "fpc_internal.hpp"

namespace fpc {

typedef unsigned char u8;

__device__ inline long long iU(int i, int j, int k, int nx, int ny) { return ((long long)k * ny + j) * (nx + 1) + i; }
__device__ inline long long iV(int i, int j, int k, int nx, int ny) { return ((long long)k * (ny + 1) + j) * nx + i; }
__device__ inline long long iW(int i, int j, int k, int nx, int ny) { return ((long long)k * ny + j) * nx + i; }
__device__ inline long long iP(int i, int j, int k, int nx, int ny) { return ((long long)k * ny + j) * nx + i; }

// Out-of-box reads return the fallback (gpu_dense_ops.hpp:1-3).
__device__ inline double gu(const double *u, int i, int j, int k, int nx, int ny, int nz, double fb) {
  return (i < 0 || i > nx || j < 0 || j >= ny || k < 0 || k >= nz) ? fb : u[iU(i, j, k, nx, ny)];
}
__device__ inline double gv(const double *v, int i, int j, int k, int nx, int ny, int nz, double fb) {
  return (i < 0 || i >= nx || j < 0 || j > ny || k < 0 || k >= nz) ? fb : v[iV(i, j, k, nx, ny)];
}
__device__ inline double gw(const double *w, int i, int j, int k, int nx, int ny, int nz, double fb) {
  return (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k > nz) ? fb : w[iW(i, j, k, nx, ny)];
}

__device__ inline double conv_u_at(const double *u, const double *v, const double *w, int i, int j, int k, int nx,
                                   int ny, int nz, double dx, double dy, double dz) {
  double uc = u[iU(i, j, k, nx, ny)];
  double va = 0.25 * (gv(v, i - 1, j, k, nx, ny, nz, 0.0) + gv(v, i - 1, j + 1, k, nx, ny, nz, 0.0) +
                      gv(v, i, j, k, nx, ny, nz, 0.0) + gv(v, i, j + 1, k, nx, ny, nz, 0.0));
  double wa = 0.25 * (gw(w, i - 1, j, k, nx, ny, nz, 0.0) + gw(w, i - 1, j, k + 1, nx, ny, nz, 0.0) +
                      gw(w, i, j, k, nx, ny, nz, 0.0) + gw(w, i, j, k + 1, nx, ny, nz, 0.0));
  double ue = gu(u, i + 1, j, k, nx, ny, nz, uc), uw = gu(u, i - 1, j, k, nx, ny, nz, uc);
  double un = gu(u, i, j + 1, k, nx, ny, nz, uc), us = gu(u, i, j - 1, k, nx, ny, nz, uc);
  double ut = gu(u, i, j, k + 1, nx, ny, nz, uc), ub = gu(u, i, j, k - 1, nx, ny, nz, uc);
  return uc * (ue - uw) / (2 * dx) + va * (un - us) / (2 * dy) + wa * (ut - ub) / (2 * dz);
}

__device__ inline double conv_v_at(const double *u, const double *v, const double *w, int i, int j, int k, int nx,
                                   int ny, int nz, double dx, double dy, double dz) {
  double vc = v[iV(i, j, k, nx, ny)];
  double ua = 0.25 * (gu(u, i, j - 1, k, nx, ny, nz, 0.0) + gu(u, i + 1, j - 1, k, nx, ny, nz, 0.0) +
                      gu(u, i, j, k, nx, ny, nz, 0.0) + gu(u, i + 1, j, k, nx, ny, nz, 0.0));
  double wa = 0.25 * (gw(w, i, j - 1, k, nx, ny, nz, 0.0) + gw(w, i, j, k, nx, ny, nz, 0.0) +
                      gw(w, i, j - 1, k + 1, nx, ny, nz, 0.0) + gw(w, i, j, k + 1, nx, ny, nz, 0.0));
  double ve = gv(v, i + 1, j, k, nx, ny, nz, vc), vw = gv(v, i - 1, j, k, nx, ny, nz, vc);
  double vn = gv(v, i, j + 1, k, nx, ny, nz, vc), vs = gv(v, i, j - 1, k, nx, ny, nz, vc);
  double vt = gv(v, i, j, k + 1, nx, ny, nz, vc), vb = gv(v, i, j, k - 1, nx, ny, nz, vc);
  return ua * (ve - vw) / (2 * dx) + vc * (vn - vs) / (2 * dy) + wa * (vt - vb) / (2 * dz);
}

__device__ inline double conv_w_at(const double *u, const double *v, const double *w, int i, int j, int k, int nx,
                                   int ny, int nz, double dx, double dy, double dz) {
  double wc = w[iW(i, j, k, nx, ny)];
  double ua = 0.25 * (gu(u, i, j, k - 1, nx, ny, nz, 0.0) + gu(u, i + 1, j, k - 1, nx, ny, nz, 0.0) +
                      gu(u, i, j, k, nx, ny, nz, 0.0) + gu(u, i + 1, j, k, nx, ny, nz, 0.0));
  double va = 0.25 * (gv(v, i, j, k - 1, nx, ny, nz, 0.0) + gv(v, i, j + 1, k - 1, nx, ny, nz, 0.0) +
                      gv(v, i, j, k, nx, ny, nz, 0.0) + gv(v, i, j + 1, k, nx, ny, nz, 0.0));
  double we = gw(w, i + 1, j, k, nx, ny, nz, wc), ww = gw(w, i - 1, j, k, nx, ny, nz, wc);
  double wn = gw(w, i, j + 1, k, nx, ny, nz, wc), ws = gw(w, i, j - 1, k, nx, ny, nz, wc);
  double wt = gw(w, i, j, k + 1, nx, ny, nz, wc), wb = gw(w, i, j, k - 1, nx, ny, nz, wc);
  return ua * (we - ww) / (2 * dx) + va * (wn - ws) / (2 * dy) + wc * (wt - wb) / (2 * dz);
}

// G p at a u face (the CSR row of immersed_operators.hpp:252-266: left then right term).
__device__ inline double gradx_at(const double *p, const u8 *fluid, int i, int j, int k, int nx, int ny, double dx) {
  double s = 0.0;
  if (i == nx) {
    long long L = iP(nx - 1, j, k, nx, ny);
    if (fluid[L]) s += (-2.0 / dx) * p[L];
    return s;
  }
  long long L = iP(i - 1, j, k, nx, ny), R = iP(i, j, k, nx, ny);
  if (fluid[L]) s += (-1.0 / dx) * p[L];
  if (fluid[R]) s += (1.0 / dx) * p[R];
  return s;
}
__device__ inline double grady_at(const double *p, const u8 *fluid, int i, int j, int k, int nx, int ny, double dy) {
  double s = 0.0;
  long long S = iP(i, j - 1, k, nx, ny), N = iP(i, j, k, nx, ny);
  if (fluid[S]) s += (-1.0 / dy) * p[S];
  if (fluid[N]) s += (1.0 / dy) * p[N];
  return s;
}
__device__ inline double gradz_at(const double *p, const u8 *fluid, int i, int j, int k, int nx, int ny, double dz) {
  double s = 0.0;
  long long B = iP(i, j, k - 1, nx, ny), T = iP(i, j, k, nx, ny);
  if (fluid[B]) s += (-1.0 / dz) * p[B];
  if (fluid[T]) s += (1.0 / dz) * p[T];
  return s;
}

// ---- the fused stages' convection and G p in two steps: every load of a face at an
// address inside its array (out-of-box neighbours read entry 0 and take the fallback
// in the evaluation), then the same expressions as conv_*_at and grad*_at on the loaded
// values -- the same bits.  Between the two the caller pins the loads so that all of
// them are in flight before the first wait.
__device__ inline double ld_u(const double *u, int i, int j, int k, int nx, int ny, int nz, bool &in) {
  in = !((i < 0) | (i > nx) | (j < 0) | (j >= ny) | (k < 0) | (k >= nz));  // no short-circuit branches
  return u[opaque(in ? iU(i, j, k, nx, ny) : 0LL)];
}
__device__ inline double ld_v(const double *v, int i, int j, int k, int nx, int ny, int nz, bool &in) {
  in = !((i < 0) | (i >= nx) | (j < 0) | (j > ny) | (k < 0) | (k >= nz));
  return v[opaque(in ? iV(i, j, k, nx, ny) : 0LL)];
}
__device__ inline double ld_w(const double *w, int i, int j, int k, int nx, int ny, int nz, bool &in) {
  in = !((i < 0) | (i >= nx) | (j < 0) | (j >= ny) | (k < 0) | (k > nz));
  return w[opaque(in ? iW(i, j, k, nx, ny) : 0LL)];
}

// Convection at a face of component comp: the centre c, the two other components' four
// values each (a, b; fallback 0) in conv_*_at's order, and the six neighbours of the same
// component (n: +x, -x, +y, -y, +z, -z; fallback c).
struct ConvLoads {
  double c, a[4], b[4], n[6];
  bool ia[4], ib[4], in[6];
};

template <int comp>
__device__ inline ConvLoads conv_loads(const double *u, const double *v, const double *w, int i, int j, int k, int nx,
                                       int ny, int nz) {
  ConvLoads L;
  bool ic;
  if (comp == 0) {
    L.c = ld_u(u, i, j, k, nx, ny, nz, ic);
    L.a[0] = ld_v(v, i - 1, j, k, nx, ny, nz, L.ia[0]);
    L.a[1] = ld_v(v, i - 1, j + 1, k, nx, ny, nz, L.ia[1]);
    L.a[2] = ld_v(v, i, j, k, nx, ny, nz, L.ia[2]);
    L.a[3] = ld_v(v, i, j + 1, k, nx, ny, nz, L.ia[3]);
    L.b[0] = ld_w(w, i - 1, j, k, nx, ny, nz, L.ib[0]);
    L.b[1] = ld_w(w, i - 1, j, k + 1, nx, ny, nz, L.ib[1]);
    L.b[2] = ld_w(w, i, j, k, nx, ny, nz, L.ib[2]);
    L.b[3] = ld_w(w, i, j, k + 1, nx, ny, nz, L.ib[3]);
    L.n[0] = ld_u(u, i + 1, j, k, nx, ny, nz, L.in[0]);
    L.n[1] = ld_u(u, i - 1, j, k, nx, ny, nz, L.in[1]);
    L.n[2] = ld_u(u, i, j + 1, k, nx, ny, nz, L.in[2]);
    L.n[3] = ld_u(u, i, j - 1, k, nx, ny, nz, L.in[3]);
    L.n[4] = ld_u(u, i, j, k + 1, nx, ny, nz, L.in[4]);
    L.n[5] = ld_u(u, i, j, k - 1, nx, ny, nz, L.in[5]);
  } else if (comp == 1) {
    L.c = ld_v(v, i, j, k, nx, ny, nz, ic);
    L.a[0] = ld_u(u, i, j - 1, k, nx, ny, nz, L.ia[0]);
    L.a[1] = ld_u(u, i + 1, j - 1, k, nx, ny, nz, L.ia[1]);
    L.a[2] = ld_u(u, i, j, k, nx, ny, nz, L.ia[2]);
    L.a[3] = ld_u(u, i + 1, j, k, nx, ny, nz, L.ia[3]);
    L.b[0] = ld_w(w, i, j - 1, k, nx, ny, nz, L.ib[0]);
    L.b[1] = ld_w(w, i, j, k, nx, ny, nz, L.ib[1]);
    L.b[2] = ld_w(w, i, j - 1, k + 1, nx, ny, nz, L.ib[2]);
    L.b[3] = ld_w(w, i, j, k + 1, nx, ny, nz, L.ib[3]);
    L.n[0] = ld_v(v, i + 1, j, k, nx, ny, nz, L.in[0]);
    L.n[1] = ld_v(v, i - 1, j, k, nx, ny, nz, L.in[1]);
    L.n[2] = ld_v(v, i, j + 1, k, nx, ny, nz, L.in[2]);
    L.n[3] = ld_v(v, i, j - 1, k, nx, ny, nz, L.in[3]);
    L.n[4] = ld_v(v, i, j, k + 1, nx, ny, nz, L.in[4]);
    L.n[5] = ld_v(v, i, j, k - 1, nx, ny, nz, L.in[5]);
  } else {
    L.c = ld_w(w, i, j, k, nx, ny, nz, ic);
    L.a[0] = ld_u(u, i, j, k - 1, nx, ny, nz, L.ia[0]);
    L.a[1] = ld_u(u, i + 1, j, k - 1, nx, ny, nz, L.ia[1]);
    L.a[2] = ld_u(u, i, j, k, nx, ny, nz, L.ia[2]);
    L.a[3] = ld_u(u, i + 1, j, k, nx, ny, nz, L.ia[3]);
    L.b[0] = ld_v(v, i, j, k - 1, nx, ny, nz, L.ib[0]);
    L.b[1] = ld_v(v, i, j + 1, k - 1, nx, ny, nz, L.ib[1]);
    L.b[2] = ld_v(v, i, j, k, nx, ny, nz, L.ib[2]);
    L.b[3] = ld_v(v, i, j + 1, k, nx, ny, nz, L.ib[3]);
    L.n[0] = ld_w(w, i + 1, j, k, nx, ny, nz, L.in[0]);
    L.n[1] = ld_w(w, i - 1, j, k, nx, ny, nz, L.in[1]);
    L.n[2] = ld_w(w, i, j + 1, k, nx, ny, nz, L.in[2]);
    L.n[3] = ld_w(w, i, j - 1, k, nx, ny, nz, L.in[3]);
    L.n[4] = ld_w(w, i, j, k + 1, nx, ny, nz, L.in[4]);
    L.n[5] = ld_w(w, i, j, k - 1, nx, ny, nz, L.in[5]);
  }
  return L;
}

__device__ inline void pin_conv(const ConvLoads &L) {
  pin_all(L.c, L.a[0], L.a[1], L.a[2], L.a[3], L.b[0], L.b[1], L.b[2], L.b[3]);
  pin_all(L.n[0], L.n[1], L.n[2], L.n[3], L.n[4], L.n[5]);
}

template <int comp>
__device__ inline double conv_eval(const ConvLoads &L, double dx, double dy, double dz) {
  const double c = L.c;
  const double A0 = L.ia[0] ? L.a[0] : 0.0, A1 = L.ia[1] ? L.a[1] : 0.0, A2 = L.ia[2] ? L.a[2] : 0.0,
               A3 = L.ia[3] ? L.a[3] : 0.0;
  const double B0 = L.ib[0] ? L.b[0] : 0.0, B1 = L.ib[1] ? L.b[1] : 0.0, B2 = L.ib[2] ? L.b[2] : 0.0,
               B3 = L.ib[3] ? L.b[3] : 0.0;
  const double e = L.in[0] ? L.n[0] : c, wv = L.in[1] ? L.n[1] : c, n = L.in[2] ? L.n[2] : c,
               s = L.in[3] ? L.n[3] : c, t = L.in[4] ? L.n[4] : c, b = L.in[5] ? L.n[5] : c;
  const double pa = 0.25 * (A0 + A1 + A2 + A3), pb = 0.25 * (B0 + B1 + B2 + B3);
  if (comp == 0) return c * (e - wv) / (2 * dx) + pa * (n - s) / (2 * dy) + pb * (t - b) / (2 * dz);
  if (comp == 1) return pa * (e - wv) / (2 * dx) + c * (n - s) / (2 * dy) + pb * (t - b) / (2 * dz);
  return pa * (e - wv) / (2 * dx) + pb * (n - s) / (2 * dy) + c * (t - b) / (2 * dz);
}

// G p at a face of component comp (grad*_at): the fluid flags and p of the cell before
// (lo) and after (hi) the face along comp, at clamped cells; a u face at i == nx has
// only its lo cell (nx - 1).
struct GradLoads {
  int flo, fhi;
  double plo, phi;
};

template <int comp>
__device__ inline GradLoads grad_loads(const double *p, const u8 *fluid, int i, int j, int k, int nx, int ny, int nz) {
  int li = i, lj = j, lk = k, hi_ = i, hj = j, hk = k;
  if (comp == 0) li = (i >= 1) ? i - 1 : 0, hi_ = (i < nx) ? i : nx - 1;
  if (comp == 1) lj = (j >= 1) ? j - 1 : 0, hj = (j < ny) ? j : ny - 1;
  if (comp == 2) lk = (k >= 1) ? k - 1 : 0, hk = (k < nz) ? k : nz - 1;
  const long long lo = iP(li, lj, lk, nx, ny), hi = iP(hi_, hj, hk, nx, ny);
  GradLoads G;
  G.flo = fluid[lo];
  G.fhi = fluid[hi];
  G.plo = p[lo];
  G.phi = p[hi];
  return G;
}

__device__ inline void pin_grad(const GradLoads &G) { pin_all(G.flo, G.fhi, G.plo, G.phi); }

template <int comp>
__device__ inline double grad_eval(const GradLoads &G, int i, int nx, double h) {
  double s = 0.0;
  if (comp == 0 && i == nx) {
    if (G.flo) s += (-2.0 / h) * G.plo;
    return s;
  }
  if (G.flo) s += (-1.0 / h) * G.plo;
  if (G.fhi) s += (1.0 / h) * G.phi;
  return s;
}

// D q at a fluid cell in CSR column order (-x, +x, -y, +y, -z, +z terms, a term skipped
// where its coefficient is zero or its face is a wall); qout supplies the outflow faces
// when non-null.  Two steps: div_loads reads every value of cell (i, j, k) at an address
// that is always in the arrays (faces i and i+1, j and j+1, k and k+1 exist for every
// cell), div_eval sums them under the tests.  Between the two the caller pins the loads
// so that all of them are in flight before the first wait.
struct DivLoads {
  double c[6], x[6], q;  // q: the outflow value of the cell's row (qout given), else 0
  bool hq;
};

__device__ inline DivLoads div_loads(const fpc_flow_ops &o, const double *u, const double *v, const double *w,
                                     const double *qout, int i, int j, int k) {
  const int nx = o.nx, ny = o.ny;
  const double *c = o.dcoef + 6 * ((long long)j * nx + i);
  DivLoads L;
#pragma unroll
  for (int q = 0; q < 6; ++q) L.c[q] = c[q];
  L.x[0] = u[iU(i, j, k, nx, ny)];
  L.x[1] = u[iU(i + 1, j, k, nx, ny)];
  L.hq = qout != nullptr;
  L.q = L.hq ? qout[(long long)j + (long long)ny * k] : 0.0;
  L.x[2] = v[iV(i, j, k, nx, ny)];
  L.x[3] = v[iV(i, j + 1, k, nx, ny)];
  L.x[4] = w[iW(i, j, k, nx, ny)];
  L.x[5] = w[iW(i, j, k + 1, nx, ny)];
  return L;
}

__device__ inline void pin_div(const DivLoads &L) {
  pin_all(L.c[0], L.c[1], L.c[2], L.c[3], L.c[4], L.c[5], L.x[0], L.x[1], L.x[2], L.x[3], L.x[4], L.x[5], L.q);
}

__device__ inline double div_eval(const DivLoads &L, int i, int j, int k, int nx, int ny, int nz) {
  double s = 0.0;
  if (i >= 1 && L.c[0] != 0.0) s += L.c[0] * L.x[0];
  if (L.c[1] != 0.0) s += L.c[1] * ((L.hq && i == nx - 1) ? L.q : L.x[1]);
  if (j >= 1 && L.c[2] != 0.0) s += L.c[2] * L.x[2];
  if (j <= ny - 2 && L.c[3] != 0.0) s += L.c[3] * L.x[3];
  if (k >= 1 && L.c[4] != 0.0) s += L.c[4] * L.x[4];
  if (k <= nz - 2 && L.c[5] != 0.0) s += L.c[5] * L.x[5];
  return s;
}

// ---------------------------------------------------------------- stage bodies
// The per-point work of the fused stages, shared by one simulation's kernels
// (dense_ops.hip) and the batched ensemble's (ensemble.hip).

// BC cleanup (gpu_dense_ops.hpp:5-10) at index tid of each of u, v, w.
__device__ inline void bc_point(long long tid, double *u, double *v, double *w, const u8 *au, const u8 *av, const u8 *aw,
                                const double *uin, int nx, int ny, int nz) {
  long long Nu = (long long)(nx + 1) * ny * nz, Nv = (long long)nx * (ny + 1) * nz, Nw = (long long)nx * ny * (nz + 1);
  if (tid < Nu) {
    long long pl = (long long)(nx + 1) * ny;
    int k = (int)(tid / pl);
    long long rem = tid - k * pl;
    int j = (int)(rem / (nx + 1));
    int i = (int)(rem - (long long)j * (nx + 1));
    if (i == 0) u[tid] = uin[j + (long long)ny * k];
    else if (i == nx) {
      long long s = iU(nx - 1, j, k, nx, ny);
      u[tid] = (nx - 1 == 0) ? uin[j + (long long)ny * k] : (au[s] ? u[s] : 0.0);
    } else if (!au[tid]) u[tid] = 0.0;
  }
  if (tid < Nv) {
    long long pl = (long long)nx * (ny + 1);
    int k = (int)(tid / pl);
    long long rem = tid - k * pl;
    int j = (int)(rem / nx);
    if (j == 0 || j == ny || !av[tid]) v[tid] = 0.0;
  }
  if (tid < Nw) {
    long long pl = (long long)nx * ny;
    int k = (int)(tid / pl);
    long long rem = tid - k * pl;
    int j = (int)(rem / nx);
    int i = (int)(rem - (long long)j * nx);
    if (k == 0 || k == nz) w[tid] = 0.0;
    else if (!aw[iW(i, j, k - 1, nx, ny)]) w[tid] = 0.0;  // the reference tests active_w[k-1]
  }
}

// Scatter of a solved box back into the full face array (gpu_dense_ops.hpp:20-22) at index tid.
__device__ inline void scatter_point(int comp, long long tid, double *o, const double *x, const double *base, int nx,
                                     int ny, int nz) {
  int sx = nx + (comp == 0), sy = ny + (comp == 1), sz = nz + (comp == 2);
  long long N = (long long)sx * sy * sz;
  if (tid >= N) return;
  if (base) o[tid] = base[tid];
  long long pl = (long long)sx * sy;
  int k = (int)(tid / pl);
  long long rem = tid - k * pl;
  int j = (int)(rem / sx);
  int i = (int)(rem - (long long)j * sx);
  if (comp == 0 && i >= 1 && i <= nx - 1) o[tid] = x[((long long)k * ny + j) * (nx - 1) + (i - 1)];
  if (comp == 1 && j >= 1 && j <= ny - 1) o[tid] = x[((long long)k * (ny - 1) + (j - 1)) * nx + i];
  if (comp == 2 && k >= 1 && k <= nz - 1) o[tid] = x[((long long)(k - 1) * ny + j) * nx + i];
}

// Momentum rhs and x0 of box row (i, j, k) of component comp (gpu_dense_ops.hpp:16-18
// with the convection and G p of the same face).
template <int comp>
__device__ inline void explicit_face(const fpc_flow_ops &o, const double *u, const double *v, const double *w,
                                     const double *p, double *rhs, double *x0, int zero, int i, int j, int k,
                                     long long row) {
  const int nx = o.nx, ny = o.ny, nz = o.nz;
  const int fi = i + (comp == 0), fj = j + (comp == 1), fk = k + (comp == 2);
  const long long t = comp == 0 ? iU(fi, fj, fk, nx, ny) : (comp == 1 ? iV(fi, fj, fk, nx, ny) : iW(fi, fj, fk, nx, ny));
  const u8 *act = comp == 0 ? o.active_u : (comp == 1 ? o.active_v : o.active_w);
  const u8 *actq = comp == 0 ? o.active_qu : (comp == 1 ? o.active_qv : o.active_qw);
  const double *scale = comp == 0 ? o.scale_u : (comp == 1 ? o.scale_v : o.scale_w);
  const double h = comp == 0 ? o.dx : (comp == 1 ? o.dy : o.dz);
  const int a = act[t], aq = actq[t];
  const int below = (comp == 2) ? act[iW(fi, fj, fk - 1, nx, ny)] : 1;  // w: convection only over an active face
  const ConvLoads C = conv_loads<comp>(u, v, w, fi, fj, fk, nx, ny, nz);
  const GradLoads G = grad_loads<comp>(p, o.fluid_p, fi, fj, fk, nx, ny, nz);
  const double sc = scale[row];
  const double bc = comp == 0 ? o.rhs_u_bc[row] : 0.0;
  pin_conv(C);
  pin_grad(G);
  pin_all(a, aq, below, sc, bc);
  const double at = o.rho / o.dt;
  const double f = C.c;
  double rr = 0.0;
  if (a) {
    const double cv = below ? conv_eval<comp>(C, o.dx, o.dy, o.dz) : 0.0;
    const double g = aq ? grad_eval<comp>(G, fi, nx, h) : 0.0;
    rr = (at * f - o.rho * cv - g + bc) * sc;
  }
  rhs[row] = rr;
  x0[row] = zero ? 0.0 : f;
}

__device__ inline void explicit_row(int comp, const fpc_flow_ops &o, const double *u, const double *v, const double *w,
                                    const double *p, double *rhs, double *x0, int zero, int i, int j, int k) {
  const int bx = o.nx - (comp == 0), by = o.ny - (comp == 1);
  const long long row = ((long long)k * by + j) * bx + i;
  if (comp == 0)
    explicit_face<0>(o, u, v, w, p, rhs, x0, zero, i, j, k, row);
  else if (comp == 1)
    explicit_face<1>(o, u, v, w, p, rhs, x0, zero, i, j, k, row);
  else
    explicit_face<2>(o, u, v, w, p, rhs, x0, zero, i, j, k, row);
}

// Pressure rhs and x0 = 0 of cell (i, j, k) (gpu_dense_ops.hpp:80-90, main.cu:314-321).
__device__ inline void divergence_cell(const fpc_flow_ops &o, const double *ut, const double *vt, const double *wt,
                                       double ramp, double *rhs, double *x0, int i, int j, int k) {
  long long c = iP(i, j, k, o.nx, o.ny);
  const int fluid = o.fluid_p[c];
  const DivLoads L = div_loads(o, ut, vt, wt, nullptr, i, j, k);
  const double bcv = o.div_bc[(long long)j + (long long)o.ny * k];
  pin_div(L);
  pin_all(fluid, bcv);
  double r = 0.0;
  if (fluid) {
    double s = div_eval(L, i, j, k, o.nx, o.ny, o.nz);
    double bc = (i == 0) ? bcv : 0.0;
    r = -(o.rho / o.dt) * (s + ramp * bc);
  }
  rhs[c] = r;
  x0[c] = 0.0;
}

// Projection of the faces at index tid of each of u, v, w (gpu_dense_ops.hpp:92-106).
__device__ inline void project_point(const fpc_flow_ops &o, long long tid, const double *phi, const double *ut,
                                     const double *vt, const double *wt, double *u, double *v, double *w, double *qout) {
  const int nx = o.nx, ny = o.ny, nz = o.nz;
  const double f = o.dt / o.rho;
  const long long Nu = (long long)(nx + 1) * ny * nz, Nv = (long long)nx * (ny + 1) * nz, Nw = (long long)nx * ny * (nz + 1);
  const bool hu = tid < Nu, hv = tid < Nv, hw = tid < Nw;
  // face coordinates (entry 0 for an index past a component's faces)
  const long long tu = hu ? tid : 0, tv = hv ? tid : 0, tw = hw ? tid : 0;
  const long long plu = (long long)(nx + 1) * ny, plv = (long long)nx * (ny + 1), plw = (long long)nx * ny;
  const int uk = (int)(tu / plu), uj = (int)((tu - uk * plu) / (nx + 1)), ui = (int)(tu - uk * plu - (long long)uj * (nx + 1));
  const int vk = (int)(tv / plv), vj = (int)((tv - vk * plv) / nx), vi = (int)(tv - vk * plv - (long long)vj * nx);
  const int wk = (int)(tw / plw), wj = (int)((tw - wk * plw) / nx), wi = (int)(tw - wk * plw - (long long)wj * nx);
  // every load of the three faces, then one wait
  const double xu = ut[tu], xv = vt[tv], xw = wt[tw];
  const int qu = o.active_qu[tu], qv = o.active_qv[tv], qw = o.active_qw[tw];
  const GradLoads Gu = grad_loads<0>(phi, o.fluid_p, ui, uj, uk, nx, ny, nz);
  const GradLoads Gv = grad_loads<1>(phi, o.fluid_p, vi, vj, vk, nx, ny, nz);
  const GradLoads Gw = grad_loads<2>(phi, o.fluid_p, wi, wj, wk, nx, ny, nz);
  pin_grad(Gu);
  pin_grad(Gv);
  pin_grad(Gw);
  pin_all(xu, xv, xw, qu, qv, qw);
  if (hu) {
    double val = xu;
    if (qu && ui >= 1) {
      val = val - f * grad_eval<0>(Gu, ui, nx, o.dx);
      if (ui == nx) qout[(long long)uj + (long long)ny * uk] = val;
    }
    u[tid] = val;
  }
  if (hv) {
    double val = xv;
    if (qv && vj >= 1 && vj <= ny - 1) val = val - f * grad_eval<1>(Gv, vi, nx, o.dy);
    v[tid] = val;
  }
  if (hw) {
    double val = xw;
    if (qw && wk >= 1 && wk <= nz - 1) val = val - f * grad_eval<2>(Gw, wi, nx, o.dz);
    w[tid] = val;
  }
}

// Blocks of the diagnostics of an nx x ny x nz box: one per 2048 entries of the longest
// face array, at most kDiagBlocks.  The partition -- and so the bits of the sums --
// depends on the box only, for one simulation and an ensemble's instances alike.
constexpr int kDiagBlocks = 2048;
__host__ __device__ inline int diag_blocks(int nx, int ny, int nz) {
  const long long Nu = (long long)(nx + 1) * ny * nz, Nv = (long long)nx * (ny + 1) * nz,
                  Nw = (long long)nx * ny * (nz + 1);
  const long long Nm = Nu > Nv ? (Nu > Nw ? Nu : Nw) : (Nv > Nw ? Nv : Nw);
  const long long b = (Nm + 2047) / 2048;
  return (int)(b < 1 ? 1 : (b > kDiagBlocks ? kDiagBlocks : b));
}

// Block bid of nblk of the diagnostics (main.cu:352-363 on the device): partial sum of the
// squared divergence, max |u|, |v|, |w|, partial forces; part[q * nblk + bid].
__device__ inline void diag_partials_body(const fpc_flow_ops &o, const double *u, const double *v, const double *w,
                                          const double *p, const double *qout, double ramp, double *part, int bid,
                                          int nblk) {
  __shared__ double lds[8];
  const int nx = o.nx, ny = o.ny, nz = o.nz;
  const long long Np = (long long)nx * ny * nz, pl = (long long)nx * ny;
  const long long Nu = (long long)(nx + 1) * ny * nz, Nv = (long long)nx * (ny + 1) * nz, Nw = pl * (nz + 1);
  const long long Nm = Nu > Nv ? (Nu > Nw ? Nu : Nw) : (Nv > Nw ? Nv : Nw);
  const long long stride = (long long)nblk * blockDim.x, t0 = bid * (long long)blockDim.x + threadIdx.x;
  // cell coordinates of t, walked by the stride's own (i, j, k)
  const int si = (int)(stride % nx), sj = (int)((stride / nx) % ny), sk = (int)(stride / pl);
  int i = (int)(t0 % nx), j = (int)((t0 / nx) % ny), k = (int)(t0 / pl);
  double d2 = 0.0, mu = 0.0, mv = 0.0, mw = 0.0, fx = 0.0, fy = 0.0;
  // One pass: cell t (t < Np) and face t of each component (max |.| is exact in any order).
  for (long long t = t0; t < Nm; t += stride) {
    const bool cell = t < Np;
    const int ci = cell ? i : 0, cj = cell ? j : 0, ck = cell ? k : 0;
    const int fluid = o.fluid_p[cell ? t : 0];
    const DivLoads L = div_loads(o, u, v, w, qout, ci, cj, ck);
    const double bcv = o.div_bc[(long long)cj + (long long)ny * ck];
    const double au = u[t < Nu ? t : 0], av = v[t < Nv ? t : 0], aw = w[t < Nw ? t : 0];
    pin_div(L);
    pin_all(fluid, bcv, au, av, aw);
    if (cell && fluid) {
      const double s = div_eval(L, ci, cj, ck, nx, ny, nz);
      const double bc = (ci == 0) ? bcv : 0.0;
      const double d = s + ramp * bc;
      d2 += d * d;
    }
    if (t < Nu) mu = fmax(mu, fabs(au));
    if (t < Nv) mv = fmax(mv, fabs(av));
    if (t < Nw) mw = fmax(mw, fabs(aw));
    ijk_advance(i, j, k, si, sj, sk, nx, ny);
  }
  for (long long t = t0; t < o.n_fu; t += stride) fx += o.fu_w[t] * u[o.fu_idx[t]];
  for (long long t = t0; t < o.n_px; t += stride) fx += o.px_w[t] * p[o.px_idx[t]];
  for (long long t = t0; t < o.n_fv; t += stride) fy += o.fv_w[t] * v[o.fv_idx[t]];
  for (long long t = t0; t < o.n_py; t += stride) fy += o.py_w[t] * p[o.py_idx[t]];
  double vals[6] = {d2, mu, mv, mw, fx, fy};
  for (int q = 0; q < 6; ++q) {
    double x = vals[q];
    bool isMax = (q >= 1 && q <= 3);
    for (int off = 32; off > 0; off >>= 1) {
      double y = __shfl_xor(x, off, 64);
      x = isMax ? fmax(x, y) : x + y;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
      double r = lds[0];
      for (int wv = 1; wv < 4; ++wv) r = isMax ? fmax(r, lds[wv]) : r + lds[wv];
      part[(long long)q * nblk + bid] = r;
    }
  }
}

// Final reduction of the nb diagnostic partials (one block of 1024): divL2 and the rest.
__device__ inline void diag_final_body(const double *part, int nb, int Np, double *out) {
  __shared__ double lds[16];
  for (int q = 0; q < 6; ++q) {
    bool isMax = (q >= 1 && q <= 3);
    double x = isMax ? 0.0 : 0.0;
    for (int t = threadIdx.x; t < nb; t += blockDim.x) x = isMax ? fmax(x, part[(long long)q * nb + t]) : x + part[(long long)q * nb + t];
    for (int off = 32; off > 0; off >>= 1) {
      double y = __shfl_xor(x, off, 64);
      x = isMax ? fmax(x, y) : x + y;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
      double r = lds[0];
      for (int wv = 1; wv < (int)(blockDim.x >> 6); ++wv) r = isMax ? fmax(r, lds[wv]) : r + lds[wv];
      if (q == 0) r = sqrt(r / (double)(Np > 1 ? Np : 1));
      out[q] = r;
    }
  }
}

}  // namespace fpc

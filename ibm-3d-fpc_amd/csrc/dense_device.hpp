// dense_device.hpp -- device helpers of the staggered-field operations (libfpc):
// full-array index maps (x fastest, common.hpp:230-238 of the reference), the
// centred convection at a face (gpu_dense_ops.hpp:1-14), G p at a face in CSR
// column order and D q at a cell in CSR column order.  Shared by the forward
// stages (dense_ops.hip) and their adjoints (adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "fpc_dense.h"

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

// D q at a fluid cell in CSR column order; qout supplies the outflow faces when non-null.
__device__ inline double div_at(const fpc_flow_ops &o, const double *u, const double *v, const double *w,
                                const double *qout, int i, int j, int k) {
  int nx = o.nx, ny = o.ny, nz = o.nz;
  const double *c = o.dcoef + 6 * ((long long)j * nx + i);
  double s = 0.0;
  if (i >= 1 && c[0] != 0.0) s += c[0] * u[iU(i, j, k, nx, ny)];
  if (c[1] != 0.0) {
    double qe = (qout && i == nx - 1) ? qout[(long long)j + (long long)ny * k] : u[iU(i + 1, j, k, nx, ny)];
    s += c[1] * qe;
  }
  if (j >= 1 && c[2] != 0.0) s += c[2] * v[iV(i, j, k, nx, ny)];
  if (j <= ny - 2 && c[3] != 0.0) s += c[3] * v[iV(i, j + 1, k, nx, ny)];
  if (k >= 1 && c[4] != 0.0) s += c[4] * w[iW(i, j, k, nx, ny)];
  if (k <= nz - 2 && c[5] != 0.0) s += c[5] * w[iW(i, j, k + 1, nx, ny)];
  return s;
}

}  // namespace fpc

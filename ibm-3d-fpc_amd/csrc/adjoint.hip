// adjoint.hip -- discrete adjoint of the time step and the checkpointed gradient
// of time-averaged drag (libfpc).  No counterpart in the reference.
//
// The forward step (main.cu:249-376, as run by sim.hip) is
//   Uc = BC(U^n)                               inlet = ramp * uin_base * theta
//   x_c = Ahat_c^-1 S_c (rho/dt Uc - rho conv_c(Uc) - G_c p^n + rhs_bc_c)   c = u, v, w
//   Ut = BC(scatter(x, Uc))
//   phi = Ap^-1 ( -(rho/dt) (D Ut + ramp div_bc_base * theta) )   on fluid cells
//   U^{n+1} = BC(Ut - (dt/rho) G phi),  p^{n+1} = p^n + phi
// with Ahat_c = S_c A_c the l1-row-scaled momentum operators.  The vector-Jacobian
// product walks these stages backwards:
//   * BC cleanup, scatter and projection are masks, copies and G: their transposes
//     are gathers (deterministic);
//   * the linear solves become solves with the transposed stencils
//     (A_c^-T = S_c Ahat_c^-T; Ap^-T with the same GMRES + PFMG-style MG built on
//     Ap^T), to the forward tolerances;
//   * the centred convection is differentiated at Uc, and its transposed
//     Jacobian is scattered with atomic adds;
//   * theta enters through every inlet assignment and the divergence inlet term;
//   * nu enters through A_c = (rho/dt) I + nu K_c and rhs_bc (both linear in nu):
//     at a converged solve dx_c/dnu = A_c^-1 ((rho/dt)(x_c - Uc) + rho conv_c + G_c p)/nu,
//     so the row scaling never needs differentiating; and through the viscous
//     drag weights (linear in nu).
// The gradient driver stores the state every k steps, recomputes each segment
// forward (keeping the pre-step state, the ramp and the momentum solutions), and
// runs the segment's adjoint steps in reverse.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "dense_device.hpp"
#include "fpc_internal.hpp"
#include "fpc_sim.h"
#include "sim_internal.hpp"

namespace fpc {
namespace {

#define ADJ_TID long long tid = blockIdx.x * (long long)blockDim.x + threadIdx.x
inline unsigned nblocks(long long n) { return (unsigned)std::max(1LL, (n + 255) / 256); }

// ------------------------------------------------------------ forward pieces
// BC cleanup (the arithmetic of the forward k_bc, gpu_dense_ops.hpp:5-10).
__global__ void k_bc_fwd(double *u, double *v, double *w, const u8 *au, const u8 *av, const u8 *aw, const double *uin,
                         int nx, int ny, int nz) {
  ADJ_TID;
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
      u[tid] = au[s] ? u[s] : 0.0;
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
    else if (!aw[iW(i, j, k - 1, nx, ny)]) w[tid] = 0.0;
  }
}

// ------------------------------------------------------------ adjoint pieces
// Transpose of BC cleanup, in place on the adjoint fields: inlet faces pass their
// adjoint to the inlet values (thb += ramp * uin_base * u_bar); the outflow copy
// hands its adjoint to the last interior face; masked faces get none.
__global__ void k_bc_adj(double *ub, double *vb, double *wb, const u8 *au, const u8 *av, const u8 *aw,
                         const double *uin_base, double ramp, double *thb, int nx, int ny, int nz) {
  ADJ_TID;
  long long Nu = (long long)(nx + 1) * ny * nz, Nv = (long long)nx * (ny + 1) * nz, Nw = (long long)nx * ny * (nz + 1);
  if (tid < Nu) {
    long long pl = (long long)(nx + 1) * ny;
    int k = (int)(tid / pl);
    long long rem = tid - k * pl;
    int j = (int)(rem / (nx + 1));
    int i = (int)(rem - (long long)j * (nx + 1));
    if (i == 0) {
      long long jk = j + (long long)ny * k;
      thb[jk] += ramp * uin_base[jk] * ub[tid];
      ub[tid] = 0.0;
    } else if (i < nx) {
      double val = au[tid] ? ub[tid] : 0.0;
      if (i == nx - 1) {
        long long o = iU(nx, j, k, nx, ny);
        if (au[tid]) val += ub[o];
        ub[o] = 0.0;
      }
      ub[tid] = val;
    }
  }
  if (tid < Nv) {
    long long pl = (long long)nx * (ny + 1);
    int k = (int)(tid / pl);
    long long rem = tid - k * pl;
    int j = (int)(rem / nx);
    if (j == 0 || j == ny || !av[tid]) vb[tid] = 0.0;
  }
  if (tid < Nw) {
    long long pl = (long long)nx * ny;
    int k = (int)(tid / pl);
    long long rem = tid - k * pl;
    int j = (int)(rem / nx);
    int i = (int)(rem - (long long)j * nx);
    if (k == 0 || k == nz || !aw[iW(i, j, k - 1, nx, ny)]) wb[tid] = 0.0;
  }
}

// G^T at a cell: the adjoint of the face values g = G p that cell c feeds (CSR
// rows of immersed_operators.hpp:252-266; the outflow face carries -2/dx).
// inc_outflow: whether the u face i = nx is one of the faces (projection: yes;
// momentum rhs: no, it is not a box row).
__device__ inline double gT_at(const fpc_flow_ops &o, const double *gu_, const double *gv_, const double *gw_, int i,
                               int j, int k, bool inc_outflow) {
  const int nx = o.nx, ny = o.ny, nz = o.nz;
  double s = 0.0;
  if (i >= 1) {
    long long f = iU(i, j, k, nx, ny);
    if (o.active_qu[f]) s += (1.0 / o.dx) * gu_[f];
  }
  {
    long long f = iU(i + 1, j, k, nx, ny);
    if (i + 1 <= nx - 1) {
      if (o.active_qu[f]) s += (-1.0 / o.dx) * gu_[f];
    } else if (inc_outflow && o.active_qu[f]) {
      s += (-2.0 / o.dx) * gu_[f];
    }
  }
  if (j >= 1) {
    long long f = iV(i, j, k, nx, ny);
    if (o.active_qv[f]) s += (1.0 / o.dy) * gv_[f];
  }
  if (j + 1 <= ny - 1) {
    long long f = iV(i, j + 1, k, nx, ny);
    if (o.active_qv[f]) s += (-1.0 / o.dy) * gv_[f];
  }
  if (k >= 1) {
    long long f = iW(i, j, k, nx, ny);
    if (o.active_qw[f]) s += (1.0 / o.dz) * gw_[f];
  }
  if (k + 1 <= nz - 1) {
    long long f = iW(i, j, k + 1, nx, ny);
    if (o.active_qw[f]) s += (-1.0 / o.dz) * gw_[f];
  }
  return s;
}

// Projection and p += phi, reversed: phib = fluid ? p_bar - (dt/rho) G^T u_bar : 0.
__global__ void k_project_adj(fpc_flow_ops o, const double *ub, const double *vb, const double *wb, const double *pb,
                              double *phib) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  if (i >= o.nx || j >= o.ny) return;
  long long c = iP(i, j, k, o.nx, o.ny);
  double r = 0.0;
  if (o.fluid_p[c]) r = pb[c] - (o.dt / o.rho) * gT_at(o, ub, vb, wb, i, j, k, true);
  phib[c] = r;
}

// Divergence + pressure rhs, reversed: u_bar += -(rho/dt) D^T (psi on fluid cells);
// the inlet term adds to thb.
__global__ void k_div_adj(fpc_flow_ops o, const double *psi, const double *divbc_base, double ramp, double *ub,
                          double *vb, double *wb, double *thb) {
  ADJ_TID;
  const int nx = o.nx, ny = o.ny, nz = o.nz;
  const double f2 = -(o.rho / o.dt);
  long long Nu = (long long)(nx + 1) * ny * nz, Nv = (long long)nx * (ny + 1) * nz, Nw = (long long)nx * ny * (nz + 1);
  auto psi_at = [&](int i, int j, int k) { long long c = iP(i, j, k, nx, ny); return o.fluid_p[c] ? psi[c] : 0.0; };
  auto cf = [&](int i, int j, int s) { return o.dcoef[6 * ((long long)j * nx + i) + s]; };
  if (tid < Nu) {
    long long pl = (long long)(nx + 1) * ny;
    int k = (int)(tid / pl);
    long long rem = tid - k * pl;
    int j = (int)(rem / (nx + 1));
    int i = (int)(rem - (long long)j * (nx + 1));
    double acc = 0.0;
    if (i <= nx - 1 && i >= 1 && cf(i, j, 0) != 0.0) acc += cf(i, j, 0) * psi_at(i, j, k);
    if (i >= 1 && cf(i - 1, j, 1) != 0.0) acc += cf(i - 1, j, 1) * psi_at(i - 1, j, k);
    ub[tid] += f2 * acc;
    if (i == 0) {
      long long jk = j + (long long)ny * k;
      if (o.fluid_p[iP(0, j, k, nx, ny)]) thb[jk] += f2 * ramp * divbc_base[jk] * psi[iP(0, j, k, nx, ny)];
    }
  }
  if (tid < Nv) {
    long long pl = (long long)nx * (ny + 1);
    int k = (int)(tid / pl);
    long long rem = tid - k * pl;
    int j = (int)(rem / nx);
    int i = (int)(rem - (long long)j * nx);
    double acc = 0.0;
    if (j <= ny - 1 && j >= 1 && cf(i, j, 2) != 0.0) acc += cf(i, j, 2) * psi_at(i, j, k);
    if (j >= 1 && j - 1 <= ny - 2 && cf(i, j - 1, 3) != 0.0) acc += cf(i, j - 1, 3) * psi_at(i, j - 1, k);
    vb[tid] += f2 * acc;
  }
  if (tid < Nw) {
    long long pl = (long long)nx * ny;
    int k = (int)(tid / pl);
    long long rem = tid - k * pl;
    int j = (int)(rem / nx);
    int i = (int)(rem - (long long)j * nx);
    double acc = 0.0;
    if (k <= nz - 1 && k >= 1 && cf(i, j, 4) != 0.0) acc += cf(i, j, 4) * psi_at(i, j, k);
    if (k >= 1 && k - 1 <= nz - 2 && cf(i, j, 5) != 0.0) acc += cf(i, j, 5) * psi_at(i, j, k - 1);
    wb[tid] += f2 * acc;
  }
}

// Scatter reversed: the box faces hand their adjoint to the solution x (0 on
// inactive rows, whose solution is identically 0) and keep none.
__global__ void k_scatter_adj(int comp, double *fb, double *xb, const u8 *act, int nx, int ny, int nz) {
  ADJ_TID;
  int bx = nx - (comp == 0), by = ny - (comp == 1), bz = nz - (comp == 2);
  long long N = (long long)bx * by * bz;
  if (tid >= N) return;
  int i = (int)(tid % bx), j = (int)((tid / bx) % by), k = (int)(tid / ((long long)bx * by));
  long long f = (comp == 0) ? iU(i + 1, j, k, nx, ny) : (comp == 1 ? iV(i, j + 1, k, nx, ny) : iW(i, j, k + 1, nx, ny));
  xb[tid] = act[f] ? fb[f] : 0.0;
  fb[f] = 0.0;
}

// Momentum rhs reversed for component comp (the arithmetic of k_explicit): with
// bb = S mu on active rows, u_bar += (rho/dt) bb, conv_bar = -rho bb, g_bar = -bb,
// and the nu partial sums bb ((rho/dt)(x - f) + rho conv + g).
__global__ void __launch_bounds__(256)
    k_rhs_adj(int comp, fpc_flow_ops o, const double *mu, const double *x, const double *u, const double *v,
              const double *w, const double *p, double *fb, double *convb, double *gb, double *nupart) {
  const int nx = o.nx, ny = o.ny, nz = o.nz;
  const int bx = nx - (comp == 0), by = ny - (comp == 1), bz = nz - (comp == 2);
  const long long N = (long long)bx * by * bz;
  const double at = o.rho / o.dt;
  double acc = 0.0;
  for (long long row = blockIdx.x * (long long)blockDim.x + threadIdx.x; row < N; row += (long long)gridDim.x * blockDim.x) {
    int i = (int)(row % bx), j = (int)((row / bx) % by), k = (int)(row / ((long long)bx * by));
    long long t;
    double f, cv = 0.0, g = 0.0, sc;
    bool act, use_conv, use_g;
    if (comp == 0) {
      t = iU(i + 1, j, k, nx, ny);
      act = o.active_u[t];
      use_conv = true;
      use_g = o.active_qu[t];
      sc = o.scale_u[row];
      if (act) {
        cv = conv_u_at(u, v, w, i + 1, j, k, nx, ny, nz, o.dx, o.dy, o.dz);
        if (use_g) g = gradx_at(p, o.fluid_p, i + 1, j, k, nx, ny, o.dx);
      }
      f = u[t];
    } else if (comp == 1) {
      t = iV(i, j + 1, k, nx, ny);
      act = o.active_v[t];
      use_conv = true;
      use_g = o.active_qv[t];
      sc = o.scale_v[row];
      if (act) {
        cv = conv_v_at(u, v, w, i, j + 1, k, nx, ny, nz, o.dx, o.dy, o.dz);
        if (use_g) g = grady_at(p, o.fluid_p, i, j + 1, k, nx, ny, o.dy);
      }
      f = v[t];
    } else {
      t = iW(i, j, k + 1, nx, ny);
      act = o.active_w[t];
      use_conv = o.active_w[iW(i, j, k, nx, ny)];  // the reference tests active_w[k-1]
      use_g = o.active_qw[t];
      sc = o.scale_w[row];
      if (act) {
        if (use_conv) cv = conv_w_at(u, v, w, i, j, k + 1, nx, ny, nz, o.dx, o.dy, o.dz);
        if (use_g) g = gradz_at(p, o.fluid_p, i, j, k + 1, nx, ny, o.dz);
      }
      f = w[t];
    }
    if (!act) {
      convb[t] = 0.0;
      gb[t] = 0.0;
      continue;
    }
    const double bb = sc * mu[row];
    fb[t] += at * bb;
    convb[t] = use_conv ? -o.rho * bb : 0.0;
    gb[t] = use_g ? -bb : 0.0;
    acc += bb * (at * (x[row] - f) + o.rho * cv + g);
  }
  __shared__ double lds[4];
  double tot = block_sum<256>(acc, lds);
  if (threadIdx.x == 0) nupart[blockIdx.x] = tot;
}

// p_bar += G^T g_bar on fluid cells (the momentum rhs uses the box faces only).
__global__ void k_grad_adj(fpc_flow_ops o, const double *gbu, const double *gbv, const double *gbw, double *pb) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  if (i >= o.nx || j >= o.ny) return;
  long long c = iP(i, j, k, o.nx, o.ny);
  if (o.fluid_p[c]) pb[c] += gT_at(o, gbu, gbv, gbw, i, j, k, false);
}

// Transposed Jacobian of the centred convection (conv_*_at), scattered with
// atomic adds into the adjoint fields; lam is the adjoint of the face's conv.
__device__ inline void add_u(double *ub, const double *u, int i, int j, int k, int nx, int ny, int nz, long long tc,
                             double val) {
  // a same-component neighbour outside the box reads as the centre value
  if (i < 0 || i > nx || j < 0 || j >= ny || k < 0 || k >= nz) atomicAdd(&ub[tc], val);
  else atomicAdd(&ub[iU(i, j, k, nx, ny)], val);
}
__device__ inline void add_v(double *vb, int i, int j, int k, int nx, int ny, int nz, long long tc, double val,
                             bool same) {
  if (i < 0 || i >= nx || j < 0 || j > ny || k < 0 || k >= nz) {
    if (same) atomicAdd(&vb[tc], val);  // cross components outside the box read as 0
  } else {
    atomicAdd(&vb[iV(i, j, k, nx, ny)], val);
  }
}
__device__ inline void add_w(double *wb, int i, int j, int k, int nx, int ny, int nz, long long tc, double val,
                             bool same) {
  if (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k > nz) {
    if (same) atomicAdd(&wb[tc], val);
  } else {
    atomicAdd(&wb[iW(i, j, k, nx, ny)], val);
  }
}

__global__ void k_conv_adj(int comp, fpc_flow_ops o, const double *u, const double *v, const double *w,
                           const double *lam_arr, double *ub, double *vb, double *wb) {
  ADJ_TID;
  const int nx = o.nx, ny = o.ny, nz = o.nz;
  const int sx = nx + (comp == 0), sy = ny + (comp == 1), sz = nz + (comp == 2);
  if (tid >= (long long)sx * sy * sz) return;
  const double lam = lam_arr[tid];
  if (lam == 0.0) return;
  int i = (int)(tid % sx), j = (int)((tid / sx) % sy), k = (int)(tid / ((long long)sx * sy));
  const double hx = 1.0 / (2 * o.dx), hy = 1.0 / (2 * o.dy), hz = 1.0 / (2 * o.dz);
  if (comp == 0) {
    const long long tc = tid;
    const double uc = u[tc];
    const double va = 0.25 * (gv(v, i - 1, j, k, nx, ny, nz, 0.0) + gv(v, i - 1, j + 1, k, nx, ny, nz, 0.0) +
                              gv(v, i, j, k, nx, ny, nz, 0.0) + gv(v, i, j + 1, k, nx, ny, nz, 0.0));
    const double wa = 0.25 * (gw(w, i - 1, j, k, nx, ny, nz, 0.0) + gw(w, i - 1, j, k + 1, nx, ny, nz, 0.0) +
                              gw(w, i, j, k, nx, ny, nz, 0.0) + gw(w, i, j, k + 1, nx, ny, nz, 0.0));
    const double ue = gu(u, i + 1, j, k, nx, ny, nz, uc), uw = gu(u, i - 1, j, k, nx, ny, nz, uc);
    const double un = gu(u, i, j + 1, k, nx, ny, nz, uc), us = gu(u, i, j - 1, k, nx, ny, nz, uc);
    const double ut = gu(u, i, j, k + 1, nx, ny, nz, uc), ubv = gu(u, i, j, k - 1, nx, ny, nz, uc);
    atomicAdd(&ub[tc], lam * (ue - uw) * hx);
    add_u(ub, u, i + 1, j, k, nx, ny, nz, tc, lam * uc * hx);
    add_u(ub, u, i - 1, j, k, nx, ny, nz, tc, -lam * uc * hx);
    add_u(ub, u, i, j + 1, k, nx, ny, nz, tc, lam * va * hy);
    add_u(ub, u, i, j - 1, k, nx, ny, nz, tc, -lam * va * hy);
    add_u(ub, u, i, j, k + 1, nx, ny, nz, tc, lam * wa * hz);
    add_u(ub, u, i, j, k - 1, nx, ny, nz, tc, -lam * wa * hz);
    const double dv = 0.25 * lam * (un - us) * hy, dw = 0.25 * lam * (ut - ubv) * hz;
    add_v(vb, i - 1, j, k, nx, ny, nz, 0, dv, false);
    add_v(vb, i - 1, j + 1, k, nx, ny, nz, 0, dv, false);
    add_v(vb, i, j, k, nx, ny, nz, 0, dv, false);
    add_v(vb, i, j + 1, k, nx, ny, nz, 0, dv, false);
    add_w(wb, i - 1, j, k, nx, ny, nz, 0, dw, false);
    add_w(wb, i - 1, j, k + 1, nx, ny, nz, 0, dw, false);
    add_w(wb, i, j, k, nx, ny, nz, 0, dw, false);
    add_w(wb, i, j, k + 1, nx, ny, nz, 0, dw, false);
  } else if (comp == 1) {
    const long long tc = tid;
    const double vc = v[tc];
    const double ua = 0.25 * (gu(u, i, j - 1, k, nx, ny, nz, 0.0) + gu(u, i + 1, j - 1, k, nx, ny, nz, 0.0) +
                              gu(u, i, j, k, nx, ny, nz, 0.0) + gu(u, i + 1, j, k, nx, ny, nz, 0.0));
    const double wa = 0.25 * (gw(w, i, j - 1, k, nx, ny, nz, 0.0) + gw(w, i, j, k, nx, ny, nz, 0.0) +
                              gw(w, i, j - 1, k + 1, nx, ny, nz, 0.0) + gw(w, i, j, k + 1, nx, ny, nz, 0.0));
    const double ve = gv(v, i + 1, j, k, nx, ny, nz, vc), vw = gv(v, i - 1, j, k, nx, ny, nz, vc);
    const double vn = gv(v, i, j + 1, k, nx, ny, nz, vc), vs = gv(v, i, j - 1, k, nx, ny, nz, vc);
    const double vt = gv(v, i, j, k + 1, nx, ny, nz, vc), vbv = gv(v, i, j, k - 1, nx, ny, nz, vc);
    atomicAdd(&vb[tc], lam * (vn - vs) * hy);
    add_v(vb, i + 1, j, k, nx, ny, nz, tc, lam * ua * hx, true);
    add_v(vb, i - 1, j, k, nx, ny, nz, tc, -lam * ua * hx, true);
    add_v(vb, i, j + 1, k, nx, ny, nz, tc, lam * vc * hy, true);
    add_v(vb, i, j - 1, k, nx, ny, nz, tc, -lam * vc * hy, true);
    add_v(vb, i, j, k + 1, nx, ny, nz, tc, lam * wa * hz, true);
    add_v(vb, i, j, k - 1, nx, ny, nz, tc, -lam * wa * hz, true);
    const double du = 0.25 * lam * (ve - vw) * hx, dw = 0.25 * lam * (vt - vbv) * hz;
    // cross-component u reads outside the box are 0: add only in-box
    if (!(i < 0 || i > nx || j - 1 < 0 || j - 1 >= ny || k < 0 || k >= nz)) atomicAdd(&ub[iU(i, j - 1, k, nx, ny)], du);
    if (!(i + 1 > nx || j - 1 < 0 || j - 1 >= ny)) atomicAdd(&ub[iU(i + 1, j - 1, k, nx, ny)], du);
    if (!(j >= ny)) atomicAdd(&ub[iU(i, j, k, nx, ny)], du);
    if (!(i + 1 > nx || j >= ny)) atomicAdd(&ub[iU(i + 1, j, k, nx, ny)], du);
    add_w(wb, i, j - 1, k, nx, ny, nz, 0, dw, false);
    add_w(wb, i, j, k, nx, ny, nz, 0, dw, false);
    add_w(wb, i, j - 1, k + 1, nx, ny, nz, 0, dw, false);
    add_w(wb, i, j, k + 1, nx, ny, nz, 0, dw, false);
  } else {
    const long long tc = tid;
    const double wc = w[tc];
    const double ua = 0.25 * (gu(u, i, j, k - 1, nx, ny, nz, 0.0) + gu(u, i + 1, j, k - 1, nx, ny, nz, 0.0) +
                              gu(u, i, j, k, nx, ny, nz, 0.0) + gu(u, i + 1, j, k, nx, ny, nz, 0.0));
    const double va = 0.25 * (gv(v, i, j, k - 1, nx, ny, nz, 0.0) + gv(v, i, j + 1, k - 1, nx, ny, nz, 0.0) +
                              gv(v, i, j, k, nx, ny, nz, 0.0) + gv(v, i, j + 1, k, nx, ny, nz, 0.0));
    const double we = gw(w, i + 1, j, k, nx, ny, nz, wc), ww = gw(w, i - 1, j, k, nx, ny, nz, wc);
    const double wn = gw(w, i, j + 1, k, nx, ny, nz, wc), ws = gw(w, i, j - 1, k, nx, ny, nz, wc);
    const double wt = gw(w, i, j, k + 1, nx, ny, nz, wc), wbv = gw(w, i, j, k - 1, nx, ny, nz, wc);
    atomicAdd(&wb[tc], lam * (wt - wbv) * hz);
    add_w(wb, i + 1, j, k, nx, ny, nz, tc, lam * ua * hx, true);
    add_w(wb, i - 1, j, k, nx, ny, nz, tc, -lam * ua * hx, true);
    add_w(wb, i, j + 1, k, nx, ny, nz, tc, lam * va * hy, true);
    add_w(wb, i, j - 1, k, nx, ny, nz, tc, -lam * va * hy, true);
    add_w(wb, i, j, k + 1, nx, ny, nz, tc, lam * wc * hz, true);
    add_w(wb, i, j, k - 1, nx, ny, nz, tc, -lam * wc * hz, true);
    const double du = 0.25 * lam * (we - ww) * hx, dv = 0.25 * lam * (wn - ws) * hy;
    if (k - 1 >= 0 && k - 1 < nz) atomicAdd(&ub[iU(i, j, k - 1, nx, ny)], du);
    if (k - 1 >= 0 && k - 1 < nz) atomicAdd(&ub[iU(i + 1, j, k - 1, nx, ny)], du);
    if (k < nz) atomicAdd(&ub[iU(i, j, k, nx, ny)], du);
    if (k < nz) atomicAdd(&ub[iU(i + 1, j, k, nx, ny)], du);
    if (k - 1 >= 0 && k - 1 < nz) atomicAdd(&vb[iV(i, j, k - 1, nx, ny)], dv);
    if (k - 1 >= 0 && k - 1 < nz) atomicAdd(&vb[iV(i, j + 1, k - 1, nx, ny)], dv);
    if (k < nz) atomicAdd(&vb[iV(i, j, k, nx, ny)], dv);
    if (k < nz) atomicAdd(&vb[iV(i, j + 1, k, nx, ny)], dv);
  }
}

// Drag functional reversed: u_bar[fu] += s fu_w, p_bar[px] += s px_w (indices may repeat).
__global__ void k_drag_adj(fpc_flow_ops o, double s, double *ub, double *pb) {
  ADJ_TID;
  if (tid < o.n_fu) atomicAdd(&ub[o.fu_idx[tid]], s * o.fu_w[tid]);
  if (tid < o.n_px) atomicAdd(&pb[o.px_idx[tid]], s * o.px_w[tid]);
}

// Viscous part of the x force: sum fu_w u[fu] (one block).
__global__ void __launch_bounds__(256) k_visc_fx(fpc_flow_ops o, const double *u, double *out) {
  double acc = 0.0;
  for (int t = threadIdx.x; t < o.n_fu; t += 256) acc += o.fu_w[t] * u[o.fu_idx[t]];
  __shared__ double lds[4];
  double tot = block_sum<256>(acc, lds);
  if (threadIdx.x == 0) *out = tot;
}

// 7-point box transpose: AT(q, opp(s)) = A(r, s) for the in-box neighbour q of r.
std::vector<double> box_transpose(const std::vector<double> &A, int bx, int by, int bz) {
  std::vector<double> T(A.size(), 0.0);
  for (int k = 0; k < bz; ++k)
    for (int j = 0; j < by; ++j)
      for (int i = 0; i < bx; ++i) {
        const long long r = ((long long)k * by + j) * bx + i;
        T[(size_t)(7 * r)] = A[(size_t)(7 * r)];
        for (int s = 1; s < 7; ++s) {
          int q[3] = {i, j, k};
          const int d = (s - 1) / 2;
          q[d] += (s & 1) ? -1 : 1;
          const int n[3] = {bx, by, bz};
          if (q[d] < 0 || q[d] >= n[d]) continue;
          const long long rq = ((long long)q[2] * by + q[1]) * bx + q[0];
          const int opp = (s & 1) ? s + 1 : s - 1;
          T[(size_t)(7 * rq + opp)] = A[(size_t)(7 * r + s)];
        }
      }
  return T;
}

}  // namespace
}  // namespace fpc

using namespace fpc;

struct fpc_adjoint_s {
  fpc_sim M = nullptr;
  fpc_StructMatrix AT[4] = {};  // u, v, w, p
  fpc_StructVector b[4] = {}, x[4] = {};
  double *pb[4] = {}, *px[4] = {};
  fpc_StructSolver s[4] = {}, pc[4] = {};
  // adjoint fields and work arrays (full staggered)
  double *ub = nullptr, *vb = nullptr, *wb = nullptr, *pbar = nullptr;
  double *uc = nullptr, *vc = nullptr, *wc = nullptr;      // BC-cleaned state
  double *cbu = nullptr, *cbv = nullptr, *cbw = nullptr;   // conv adjoints
  double *gbu = nullptr, *gbv = nullptr, *gbw = nullptr;   // G p adjoints
  double *phib = nullptr, *thb = nullptr, *uinb = nullptr, *dbase = nullptr, *ubase = nullptr;
  double *nupart = nullptr, *visc = nullptr;
  double *hred = nullptr;  // pinned
  long long Nu = 0, Nv = 0, Nw = 0, Np = 0, Bn[3] = {0, 0, 0};
  // tape: per step of the current segment
  struct Entry {
    double *u = nullptr, *v = nullptr, *w = nullptr, *p = nullptr, *xu = nullptr, *xv = nullptr, *xw = nullptr;
    double ramp = 0.0;
  };
  std::vector<Entry> tape;
  std::vector<double *> ckpt;  // 4 fields per checkpoint, packed
  double last_J = 0.0;
};

namespace {

__global__ void k_scale(const double *a, double s, double *out, int n) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = s * a[t];
}

int dalloc(double **p, long long n) {
  FPC_HIP(hipMalloc((void **)p, (size_t)std::max(1LL, n) * sizeof(double)));
  FPC_HIP(hipMemsetAsync(*p, 0, (size_t)std::max(1LL, n) * sizeof(double), stream()));
  return 0;
}

int dcopy(double *dst, const double *src, long long n) {
  FPC_HIP(hipMemcpyAsync(dst, src, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, stream()));
  return 0;
}

int dzero(double *p, long long n) {
  FPC_HIP(hipMemsetAsync(p, 0, (size_t)n * sizeof(double), stream()));
  return 0;
}

int setup_solver(fpc_StructSolver *s, fpc_StructSolver *pc, const fpc_params &par, int which, fpc_StructMatrix A,
                 fpc_StructVector b, fpc_StructVector x) {
  int rc = fpc_StructGMRESCreate(s);
  if (which < 3) {
    rc |= fpc_StructGMRESSetTol(*s, par.vel_tol);
    rc |= fpc_StructGMRESSetMaxIter(*s, par.vel_maxit);
    rc |= fpc_StructGMRESSetKDim(*s, par.vel_kdim);
    rc |= fpc_StructJacobiCreate(pc);
    rc |= fpc_StructJacobiSetMaxIter(*pc, par.vel_jacobi_iters);
    rc |= fpc_StructJacobiSetZeroGuess(*pc);
    rc |= fpc_StructGMRESSetPrecond(*s, fpc_StructJacobiSolve, fpc_StructJacobiSetup, *pc);
  } else {
    rc |= fpc_StructGMRESSetTol(*s, par.p_tol);
    rc |= fpc_StructGMRESSetMaxIter(*s, par.p_maxit);
    rc |= fpc_StructGMRESSetKDim(*s, par.p_kdim);
    rc |= fpc_StructPFMGCreate(pc);
    rc |= fpc_StructPFMGSetMaxIter(*pc, par.p_pc_maxit);
    rc |= fpc_StructPFMGSetTol(*pc, 0.0);
    rc |= fpc_StructPFMGSetRelaxType(*pc, par.p_pc_relax_type);
    rc |= fpc_StructPFMGSetRAPType(*pc, par.p_pc_rap_type);
    rc |= fpc_StructPFMGSetZeroGuess(*pc);
    // The transposed solves start from zero every step; a 64-point coarsest
    // level takes less time over them than the size-scaled default.
    rc |= fpc_StructPFMGSetCoarseSize(*pc, 64);
    rc |= fpc_StructGMRESSetPrecond(*s, fpc_StructPFMGSolve, fpc_StructPFMGSetup, *pc);
  }
  if (rc) return rc;
  return fpc_StructGMRESSetup(*s, A, b, x);
}

// Transposed solve `which` (0..2 momentum, 3 pressure) from a zero start.
int solve_t(fpc_adjoint A, int which, long long n) {
  int rc = dzero(A->px[which], n);
  if (rc) return rc;
  rc = fpc_StructGMRESSolve(A->s[which], A->AT[which], A->b[which], A->x[which]);
  if (rc & FPC_ERROR_CONV) fpc_ClearError(FPC_ERROR_CONV);
  return rc & ~FPC_ERROR_CONV;
}

// Reverse of one step.  e: the step's tape entry (state before the step, its
// momentum solutions, its ramp).  Turns the adjoint of the state after the step
// (ub, vb, wb, pbar) into the adjoint of the state before it, adds the inlet
// adjoint to thb and the momentum nu term to *nu_acc (not yet divided by nu).
int adjoint_step(fpc_adjoint A, const fpc_adjoint_s::Entry &e, double *nu_acc) {
  fpc_sim M = A->M;
  const fpc_flow_ops &o = M->ops;
  const int nx = o.nx, ny = o.ny, nz = o.nz;
  hipStream_t st = stream();
  const long long Nmax = std::max(A->Nu, std::max(A->Nv, A->Nw));
  Dims dp;
  dp.nx = nx; dp.ny = ny; dp.nz = nz; dp.N = A->Np;
  dim3 gp, bp;
  launch_shape(dp, gp, bp);
  int rc;
  // phase 6 reversed: BC cleanup, projection, p += phi  ->  rhs of Ap^T psi = phib
  k_bc_adj<<<nblocks(Nmax), 256, 0, st>>>(A->ub, A->vb, A->wb, o.active_u, o.active_v, o.active_w, A->ubase, e.ramp,
                                          A->thb, nx, ny, nz);
  FPC_LAUNCH_CHECK();
  k_project_adj<<<gp, bp, 0, st>>>(o, A->ub, A->vb, A->wb, A->pbar, A->pb[3]);
  FPC_LAUNCH_CHECK();
  // phase 5 reversed
  if ((rc = solve_t(A, 3, A->Np))) return rc;
  // phase 4 reversed: divergence (and its inlet term)
  k_div_adj<<<nblocks(Nmax), 256, 0, st>>>(o, A->px[3], A->dbase, e.ramp, A->ub, A->vb, A->wb, A->thb);
  FPC_LAUNCH_CHECK();
  // phase 3 reversed: BC cleanup, then the scatter hands the box faces to x
  k_bc_adj<<<nblocks(Nmax), 256, 0, st>>>(A->ub, A->vb, A->wb, o.active_u, o.active_v, o.active_w, A->ubase, e.ramp,
                                          A->thb, nx, ny, nz);
  FPC_LAUNCH_CHECK();
  double *fb[3] = {A->ub, A->vb, A->wb};
  const u8 *act[3] = {o.active_u, o.active_v, o.active_w};
  for (int c = 0; c < 3; ++c) {
    k_scatter_adj<<<nblocks(A->Bn[c]), 256, 0, st>>>(c, fb[c], A->pb[c], act[c], nx, ny, nz);
    FPC_LAUNCH_CHECK();
  }
  // phase 2 reversed: mu_c = Ahat_c^-T x_bar_c
  for (int c = 0; c < 3; ++c)
    if ((rc = solve_t(A, c, A->Bn[c]))) return rc;
  // the cleaned state of this step, recomputed
  const int nin = ny * nz;
  k_scale<<<(nin + 255) / 256, 256, 0, st>>>(M->uin, e.ramp, A->uinb, nin);
  FPC_LAUNCH_CHECK();
  if ((rc = dcopy(A->uc, e.u, A->Nu)) || (rc = dcopy(A->vc, e.v, A->Nv)) || (rc = dcopy(A->wc, e.w, A->Nw))) return rc;
  k_bc_fwd<<<nblocks(Nmax), 256, 0, st>>>(A->uc, A->vc, A->wc, o.active_u, o.active_v, o.active_w, A->uinb, nx, ny,
                                          nz);
  FPC_LAUNCH_CHECK();
  // phase 1 reversed: momentum rhs, G p, convection, BC cleanup
  double *cb[3] = {A->cbu, A->cbv, A->cbw}, *gb[3] = {A->gbu, A->gbv, A->gbw};
  const double *xs[3] = {e.xu, e.xv, e.xw};
  for (int c = 0; c < 3; ++c) {
    k_rhs_adj<<<kRedBlocks, 256, 0, st>>>(c, o, A->px[c], xs[c], A->uc, A->vc, A->wc, e.p, fb[c], cb[c], gb[c],
                                         A->nupart);
    FPC_LAUNCH_CHECK();
    if ((rc = reduce_partials(A->nupart, kRedBlocks, A->visc + 1 + c))) return rc;
  }
  k_grad_adj<<<gp, bp, 0, st>>>(o, A->gbu, A->gbv, A->gbw, A->pbar);
  FPC_LAUNCH_CHECK();
  const long long nf[3] = {A->Nu, A->Nv, A->Nw};
  for (int c = 0; c < 3; ++c) {
    k_conv_adj<<<nblocks(nf[c]), 256, 0, st>>>(c, o, A->uc, A->vc, A->wc, cb[c], A->ub, A->vb, A->wb);
    FPC_LAUNCH_CHECK();
  }
  k_bc_adj<<<nblocks(Nmax), 256, 0, st>>>(A->ub, A->vb, A->wb, o.active_u, o.active_v, o.active_w, A->ubase, e.ramp,
                                          A->thb, nx, ny, nz);
  FPC_LAUNCH_CHECK();
  FPC_HIP(hipMemcpyAsync(A->hred, A->visc + 1, 3 * sizeof(double), hipMemcpyDeviceToHost, st));
  FPC_HIP(hipStreamSynchronize(st));
  *nu_acc += A->hred[0] + A->hred[1] + A->hred[2];
  return 0;
}

// Forward phases 0-2 of the step that starts from the simulation's current
// state, keeping what the reverse needs in tape entry e; then phases 3-6.
int record_step(fpc_adjoint A, fpc_adjoint_s::Entry &e, bool finish) {
  fpc_sim M = A->M;
  int rc;
  if ((rc = dcopy(e.u, M->u, A->Nu)) || (rc = dcopy(e.v, M->v, A->Nv)) || (rc = dcopy(e.w, M->w, A->Nw)) ||
      (rc = dcopy(e.p, M->p, A->Np)))
    return rc;
  for (int ph = 0; ph <= 2; ++ph)
    if ((rc = sim_run_phase(M, ph))) return rc;
  e.ramp = M->ramp;
  if ((rc = dcopy(e.xu, M->pxu, A->Bn[0])) || (rc = dcopy(e.xv, M->pxv, A->Bn[1])) || (rc = dcopy(e.xw, M->pxw, A->Bn[2])))
    return rc;
  if (finish)
    for (int ph = 3; ph <= 6; ++ph)
      if ((rc = sim_run_phase(M, ph))) return rc;
  return 0;
}

int ensure_tape(fpc_adjoint A, int n) {
  while ((int)A->tape.size() < n) {
    fpc_adjoint_s::Entry e;
    int rc = 0;
    rc |= dalloc(&e.u, A->Nu);
    rc |= dalloc(&e.v, A->Nv);
    rc |= dalloc(&e.w, A->Nw);
    rc |= dalloc(&e.p, A->Np);
    rc |= dalloc(&e.xu, A->Bn[0]);
    rc |= dalloc(&e.xv, A->Bn[1]);
    rc |= dalloc(&e.xw, A->Bn[2]);
    A->tape.push_back(e);
    if (rc) return rc;
  }
  return 0;
}

int state_save(fpc_adjoint A, double *dst) {
  fpc_sim M = A->M;
  int rc;
  if ((rc = dcopy(dst, M->u, A->Nu)) || (rc = dcopy(dst + A->Nu, M->v, A->Nv)) ||
      (rc = dcopy(dst + A->Nu + A->Nv, M->w, A->Nw)) || (rc = dcopy(dst + A->Nu + A->Nv + A->Nw, M->p, A->Np)))
    return rc;
  return 0;
}

int state_load(fpc_adjoint A, const double *src) {
  fpc_sim M = A->M;
  int rc;
  if ((rc = dcopy(M->u, src, A->Nu)) || (rc = dcopy(M->v, src + A->Nu, A->Nv)) ||
      (rc = dcopy(M->w, src + A->Nu + A->Nv, A->Nw)) || (rc = dcopy(M->p, src + A->Nu + A->Nv + A->Nw, A->Np)))
    return rc;
  return 0;
}

double drag_den(const fpc_params &P) {
  const double ub = 4.0 * P.Um / 9.0;
  return P.rho * ub * ub * (2.0 * P.Rcirc) * P.Lz;
}

}  // namespace

extern "C" {

int fpc_adjoint_create(fpc_sim M, fpc_adjoint *out) {
  FPC_CHECK_ARG(M && out, "null argument");
  StreamScope scope(M->stream);
  fpc_adjoint A = new fpc_adjoint_s();
  A->M = M;
  const fpc_params &P = M->par;
  const int nx = P.nx, ny = P.ny, nz = P.nz;
  A->Nu = (long long)(nx + 1) * ny * nz;
  A->Nv = (long long)nx * (ny + 1) * nz;
  A->Nw = (long long)nx * ny * (nz + 1);
  A->Np = (long long)nx * ny * nz;
  A->Bn[0] = (long long)(nx - 1) * ny * nz;
  A->Bn[1] = (long long)nx * (ny - 1) * nz;
  A->Bn[2] = (long long)nx * ny * (nz - 1);
  const int bd[4][3] = {{nx - 1, ny, nz}, {nx, ny - 1, nz}, {nx, ny, nz - 1}, {nx, ny, nz}};
  const std::vector<double> *vals[4] = {&M->S.Au, &M->S.Av, &M->S.Aw, &M->S.Ap};
  fpc_StructGrid grids[4] = {M->gu, M->gv, M->gw, M->gp};
  int rc = 0;
  for (int c = 0; c < 4 && !rc; ++c) {
    std::vector<double> T = box_transpose(*vals[c], bd[c][0], bd[c][1], bd[c][2]);
    A->AT[c] = fpc_CreateMatrix3D(grids[c], M->st, bd[c][0], bd[c][1], bd[c][2], T.data());
    A->b[c] = fpc_CreateVector3D(grids[c]);
    A->x[c] = fpc_CreateVector3D(grids[c]);
    if (!A->AT[c] || !A->b[c] || !A->x[c]) {
      rc = FPC_ERROR_GENERIC;
      break;
    }
    fpc_StructVectorGetDataPointer(A->b[c], &A->pb[c]);
    fpc_StructVectorGetDataPointer(A->x[c], &A->px[c]);
    rc = setup_solver(&A->s[c], &A->pc[c], P, c, A->AT[c], A->b[c], A->x[c]);
  }
  const long long nin = (long long)ny * nz;
  rc |= dalloc(&A->ub, A->Nu) | dalloc(&A->vb, A->Nv) | dalloc(&A->wb, A->Nw) | dalloc(&A->pbar, A->Np);
  rc |= dalloc(&A->uc, A->Nu) | dalloc(&A->vc, A->Nv) | dalloc(&A->wc, A->Nw);
  rc |= dalloc(&A->cbu, A->Nu) | dalloc(&A->cbv, A->Nv) | dalloc(&A->cbw, A->Nw);
  rc |= dalloc(&A->gbu, A->Nu) | dalloc(&A->gbv, A->Nv) | dalloc(&A->gbw, A->Nw);
  rc |= dalloc(&A->thb, nin) | dalloc(&A->uinb, nin) | dalloc(&A->dbase, nin) | dalloc(&A->ubase, nin);
  rc |= dalloc(&A->nupart, kRedBlocks) | dalloc(&A->visc, 8);
  if (!rc && hipHostMalloc((void **)&A->hred, 8 * sizeof(double)) != hipSuccess) rc = FPC_ERROR_MEMORY;
  if (!rc) {
    FPC_HIP(hipMemcpyAsync(A->ubase, M->uin_base.data(), (size_t)nin * sizeof(double), hipMemcpyHostToDevice, stream()));
    FPC_HIP(hipMemcpyAsync(A->dbase, M->divbc_base.data(), (size_t)nin * sizeof(double), hipMemcpyHostToDevice,
                           stream()));
    FPC_HIP(hipStreamSynchronize(stream()));
  }
  if (rc) {
    fpc_adjoint_destroy(A);
    return rc;
  }
  *out = A;
  return 0;
}

int fpc_adjoint_step_vjp(fpc_adjoint A, int step, const double *u, const double *v, const double *w,
                         const double *p, double *au, double *av, double *aw, double *ap, double *theta_bar,
                         double *nu_bar) {
  FPC_CHECK_ARG(A && u && v && w && p && au && av && aw && ap && theta_bar && nu_bar, "null argument");
  FPC_CHECK_ARG(step >= 1, "step is the 1-based index of the step");
  fpc_sim M = A->M;
  StreamScope scope(M->stream);
  int rc = ensure_tape(A, 1);
  if (rc) return rc;
  fpc_adjoint_s::Entry &e = A->tape[0];
  // run the forward step's first phases from the given state
  if ((rc = dcopy(M->u, u, A->Nu)) || (rc = dcopy(M->v, v, A->Nv)) || (rc = dcopy(M->w, w, A->Nw)) ||
      (rc = dcopy(M->p, p, A->Np)))
    return rc;
  M->step = step - 1;
  if ((rc = record_step(A, e, false))) return rc;
  if ((rc = dcopy(A->ub, au, A->Nu)) || (rc = dcopy(A->vb, av, A->Nv)) || (rc = dcopy(A->wb, aw, A->Nw)) ||
      (rc = dcopy(A->pbar, ap, A->Np)))
    return rc;
  const long long nin = (long long)M->par.ny * M->par.nz;
  if ((rc = dcopy(A->thb, theta_bar, nin))) return rc;
  double nacc = 0.0;
  if ((rc = adjoint_step(A, e, &nacc))) return rc;
  if ((rc = dcopy(au, A->ub, A->Nu)) || (rc = dcopy(av, A->vb, A->Nv)) || (rc = dcopy(aw, A->wb, A->Nw)) ||
      (rc = dcopy(ap, A->pbar, A->Np)) || (rc = dcopy(theta_bar, A->thb, nin)))
    return rc;
  FPC_HIP(hipStreamSynchronize(stream()));
  *nu_bar += nacc / M->par.nu;
  return 0;
}

int fpc_adjoint_drag_gradient(fpc_adjoint A, int nsteps, int avg_from, int every, double *J, double *grad_theta,
                              double *grad_nu) {
  FPC_CHECK_ARG(A && J && grad_theta && grad_nu, "null argument");
  FPC_CHECK_ARG(nsteps >= 1 && every >= 1 && avg_from >= 1 && avg_from <= nsteps, "bad step window");
  fpc_sim M = A->M;
  StreamScope scope(M->stream);
  const fpc_params &P = M->par;
  const long long S4 = A->Nu + A->Nv + A->Nw + A->Np, nin = (long long)P.ny * P.nz;
  const int nseg = (nsteps + every - 1) / every;
  int rc = 0;
  while ((int)A->ckpt.size() < nseg + 1) {
    double *c = nullptr;
    if ((rc = dalloc(&c, S4))) return rc;
    A->ckpt.push_back(c);
  }
  if ((rc = ensure_tape(A, every))) return rc;
  const int step0 = M->step;
  const double den = drag_den(P);
  const int window = nsteps - avg_from + 1;
  const double sc = 2.0 / (den * window);  // dJ/dFx at a step of the window
  // forward sweep: checkpoints, the objective and the viscous part of the drag
  double Jsum = 0.0, visc_sum = 0.0;
  for (int n = 0; n < nsteps; ++n) {
    if (n % every == 0 && (rc = state_save(A, A->ckpt[(size_t)(n / every)]))) return rc;
    for (int ph = 0; ph <= 7; ++ph)
      if ((rc = sim_run_phase(M, ph))) return rc;
    if (n + 1 >= avg_from) {
      Jsum += 2.0 * M->hdiag[4] / den;
      k_visc_fx<<<1, 256, 0, stream()>>>(M->ops, M->u, A->visc);
      FPC_LAUNCH_CHECK();
      FPC_HIP(hipMemcpyAsync(A->hred, A->visc, sizeof(double), hipMemcpyDeviceToHost, stream()));
      FPC_HIP(hipStreamSynchronize(stream()));
      visc_sum += A->hred[0];
    }
  }
  if ((rc = state_save(A, A->ckpt[(size_t)nseg]))) return rc;
  // reverse sweep, segment by segment
  if ((rc = dzero(A->ub, A->Nu)) || (rc = dzero(A->vb, A->Nv)) || (rc = dzero(A->wb, A->Nw)) ||
      (rc = dzero(A->pbar, A->Np)) || (rc = dzero(A->thb, nin)))
    return rc;
  double nacc = 0.0;
  const long long nf = std::max((long long)M->ops.n_fu, (long long)M->ops.n_px);
  for (int sgm = nseg - 1; sgm >= 0; --sgm) {
    if ((rc = state_load(A, A->ckpt[(size_t)sgm]))) return rc;
    M->step = step0 + sgm * every;
    const int ns = std::min(every, nsteps - sgm * every);
    for (int q = 0; q < ns; ++q)
      if ((rc = record_step(A, A->tape[(size_t)q], true))) return rc;
    for (int q = ns - 1; q >= 0; --q) {
      if (sgm * every + q + 1 >= avg_from) {
        k_drag_adj<<<nblocks(nf), 256, 0, stream()>>>(M->ops, sc, A->ub, A->pbar);
        FPC_LAUNCH_CHECK();
      }
      if ((rc = adjoint_step(A, A->tape[(size_t)q], &nacc))) return rc;
    }
  }
  // restore the state at the end of the run
  if ((rc = state_load(A, A->ckpt[(size_t)nseg]))) return rc;
  M->step = step0 + nsteps;
  FPC_HIP(hipMemcpyAsync(grad_theta, A->thb, (size_t)nin * sizeof(double), hipMemcpyDeviceToHost, stream()));
  FPC_HIP(hipStreamSynchronize(stream()));
  *J = Jsum / window;
  *grad_nu = (nacc + sc * visc_sum) / P.nu;
  A->last_J = *J;
  return 0;
}

int fpc_adjoint_destroy(fpc_adjoint A) {
  if (!A) return 0;
  StreamScope scope(A->M->stream);
  (void)hipStreamSynchronize(stream());
  for (int c = 0; c < 4; ++c) {
    if (A->s[c]) fpc_StructGMRESDestroy(A->s[c]);
    if (A->pc[c]) (c < 3) ? fpc_StructJacobiDestroy(A->pc[c]) : fpc_StructPFMGDestroy(A->pc[c]);
    if (A->b[c]) fpc_StructVectorDestroy(A->b[c]);
    if (A->x[c]) fpc_StructVectorDestroy(A->x[c]);
    if (A->AT[c]) fpc_StructMatrixDestroy(A->AT[c]);
  }
  double *bufs[] = {A->ub, A->vb, A->wb, A->pbar, A->uc, A->vc, A->wc, A->cbu, A->cbv, A->cbw, A->gbu, A->gbv,
                    A->gbw, A->phib, A->thb, A->uinb, A->dbase, A->ubase, A->nupart, A->visc};
  for (double *b : bufs)
    if (b) (void)hipFree(b);
  for (auto &e : A->tape)
    for (double *b : {e.u, e.v, e.w, e.p, e.xu, e.xv, e.xw})
      if (b) (void)hipFree(b);
  for (double *c : A->ckpt)
    if (c) (void)hipFree(c);
  if (A->hred) (void)hipHostFree(A->hred);
  delete A;
  return 0;
}

}  // extern "C"

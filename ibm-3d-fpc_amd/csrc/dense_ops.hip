// dense_ops.hip -- per-timestep dense operations on the staggered fields (libfpc).
//
// Replaces gpu_dense_ops.hpp of the reference (JD63021/ibm-3d-fpc).  Layer 1
// restates each reference kernel with the same argument meaning; layer 2 fuses
// the pack -> CSR -> unpack sequences of main.cu:262-350 into structured MAC
// stencils.  The structured D and G sum their terms in the CSR column order
// (west, east, south, north, bottom, top for D; left then right for G), so the
// results equal the CSR path's.
//
// Arithmetic note: device code is compiled with FP contraction, so a*b + c
// becomes one fused multiply-add, as nvcc does for the reference kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "fpc_dense.h"
#include "dense_device.hpp"
#include "fpc_internal.hpp"

// Partial sums of fpc_stage_diagnostics_noalloc: kDiagBlocks blocks x 6 sums (96 KiB),
// one copy per GPU, held until the process ends.  In host code this name is a host-side
// stand-in, not a device address: the entry point looks the address up per device
// (hipGetSymbolAddress) and passes that to the kernels.
static __device__ __align__(256) double scratch[fpc::kDiagBlocks * 6];

namespace fpc {
namespace {

#define TID_1D long long tid = blockIdx.x * (long long)blockDim.x + threadIdx.x

// ---------------------------------------------------------------- layer 1 kernels
__global__ void k_bc(double *u, double *v, double *w, const u8 *au, const u8 *av, const u8 *aw, const double *uin,
                     int nx, int ny, int nz) {
  TID_1D;
  bc_point(tid, u, v, w, au, av, aw, uin, nx, ny, nz);
}

__global__ void k_conv(int comp, double *conv, const double *u, const double *v, const double *w, const u8 *act, int nx,
                       int ny, int nz, double dx, double dy, double dz) {
  TID_1D;
  int sx = nx + (comp == 0), sy = ny + (comp == 1), sz = nz + (comp == 2);
  long long N = (long long)sx * sy * sz;
  if (tid >= N) return;
  long long pl = (long long)sx * sy;
  int k = (int)(tid / pl);
  long long rem = tid - k * pl;
  int j = (int)(rem / sx);
  int i = (int)(rem - (long long)j * sx);
  conv[tid] = 0.0;
  if (comp == 0) {
    if (i < 1 || i > nx - 1 || !act[tid]) return;
    conv[tid] = conv_u_at(u, v, w, i, j, k, nx, ny, nz, dx, dy, dz);
  } else if (comp == 1) {
    if (j < 1 || j > ny - 1 || !act[tid]) return;
    conv[tid] = conv_v_at(u, v, w, i, j, k, nx, ny, nz, dx, dy, dz);
  } else {
    if (k < 1 || k > nz - 1 || !act[iW(i, j, k - 1, nx, ny)]) return;
    conv[tid] = conv_w_at(u, v, w, i, j, k, nx, ny, nz, dx, dy, dz);
  }
}

__global__ void k_build_full(int comp, double *rhsf, double *x0, const double *f, const double *conv,
                             const double *gp, const u8 *act, const int *id, const double *bc, const double *scale,
                             int nx, int ny, int nz, double rho, double dt, int zero) {
  TID_1D;
  int bx = nx - (comp == 0), by = ny - (comp == 1), bz = nz - (comp == 2);
  long long N = (long long)bx * by * bz;
  if (tid >= N) return;
  long long pl = (long long)bx * by;
  int k = (int)(tid / pl);
  long long rem = tid - k * pl;
  int j = (int)(rem / bx);
  int i = (int)(rem - (long long)j * bx);
  long long idx = comp == 0 ? iU(i + 1, j, k, nx, ny) : (comp == 1 ? iV(i, j + 1, k, nx, ny) : iW(i, j, k + 1, nx, ny));
  int aid = id[idx];
  double r = 0.0;
  if (aid > 0 && act[idx]) r = ((rho / dt) * f[idx] - rho * conv[idx] - gp[idx] + bc[aid]) * scale[aid];
  rhsf[tid] = r;
  x0[tid] = zero ? 0.0 : f[idx];
}

__global__ void k_scatter(int comp, double *o, const double *x, const double *base, int nx, int ny, int nz) {
  TID_1D;
  scatter_point(comp, tid, o, x, base, nx, ny, nz);
}

__global__ void k_csr(int nrows, const int *rowptr, const int *colind, const double *vals, const double *x, double *y) {
  TID_1D;
  if (tid >= nrows) return;
  double s = 0.0;
  for (int jj = rowptr[tid]; jj < rowptr[tid + 1]; ++jj) s += vals[jj] * x[colind[jj]];
  y[tid + 1] = s;
}

__global__ void k_pack_masked(double *vec, const double *p, const u8 *mask, const int *id, long long N) {
  TID_1D;
  if (tid >= N) return;
  if (mask[tid]) {
    int a = id[tid];
    if (a > 0) vec[a] = p[tid];
  }
}

__global__ void k_unpack_gp(const double *gp, double *out, const int *id, long long N, long long off) {
  TID_1D;
  if (tid >= N) return;
  int a = id[tid];
  out[tid] = a > 0 ? gp[off + a] : 0.0;
}

__global__ void k_pack_q(double *q, const double *f, const int *id, long long N, long long off) {
  TID_1D;
  if (tid >= N) return;
  int a = id[tid];
  if (a > 0) q[off + a] = f[tid];
}

__global__ void k_unpack_q(const double *q, double *f, const int *id, long long N, long long off) {
  TID_1D;
  if (tid >= N) return;
  int a = id[tid];
  if (a > 0) f[tid] = q[off + a];
}

__global__ void k_prhs(double *rhsf, double *x0, const double *Dqt, const double *dbc, const u8 *fluid, const int *id,
                       long long N, double rho, double dt, double ramp) {
  TID_1D;
  if (tid >= N) return;
  double r = 0.0;
  if (fluid[tid]) {
    int a = id[tid];
    if (a > 0) r = -(rho / dt) * (Dqt[a] + ramp * dbc[a]);
  }
  rhsf[tid] = r;
  x0[tid] = 0.0;
}

__global__ void k_qcorr(double *q, const double *g, double f, long long n1) {
  TID_1D;
  if (tid >= n1) return;
  if (tid > 0) q[tid] -= f * g[tid];
}

__global__ void k_add_phi(double *p, const double *phi, const u8 *fluid, const int *id, long long N) {
  TID_1D;
  if (tid >= N) return;
  if (fluid[tid]) {
    int a = id[tid];
    if (a > 0) p[tid] += phi[a];
  }
}

__global__ void k_zero(double *p, long long n) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) p[t] = 0.0;
}

// ---------------------------------------------------------------- layer 2 kernels
// Momentum rhs and x0 on a u/v/w box row, one thread per row.
__global__ void k_explicit(int comp, fpc_flow_ops o, const double *u, const double *v, const double *w, const double *p,
                           double *rhs, double *x0, int zero) {
  int bx = o.nx - (comp == 0), by = o.ny - (comp == 1), bz = o.nz - (comp == 2);
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  if (i >= bx || j >= by || k >= bz) return;
  explicit_row(comp, o, u, v, w, p, rhs, x0, zero, i, j, k);
}

// Pressure rhs on every cell of the p box.
__global__ void k_divergence(fpc_flow_ops o, const double *ut, const double *vt, const double *wt, double ramp,
                             double *rhs, double *x0) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  if (i >= o.nx || j >= o.ny) return;
  divergence_cell(o, ut, vt, wt, ramp, rhs, x0, i, j, k);
}

// Projection of the face velocities; thread t covers face t of each of u, v, w.
__global__ void k_project(fpc_flow_ops o, const double *phi, const double *ut, const double *vt, const double *wt,
                          double *u, double *v, double *w, double *qout) {
  TID_1D;
  project_point(o, tid, phi, ut, vt, wt, u, v, w, qout);
}

__global__ void k_add_phi_box(double *p, const double *phi, const u8 *fluid, long long N) {
  TID_1D;
  if (tid >= N) return;
  if (fluid[tid]) p[tid] += phi[tid];
}

// Partial diagnostics per block: sum of squared divergence, max |u|,|v|,|w|, forces.
__global__ void __launch_bounds__(256) k_diag_partials(fpc_flow_ops o, const double *u, const double *v, const double *w,
                                                       const double *p, const double *qout, double ramp,
                                                       double *part) {
  diag_partials_body(o, u, v, w, p, qout, ramp, part, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(1024) k_diag_final(const double *part, int nb, int Np, double *out) {
  diag_final_body(part, nb, Np, out);
}

__global__ void k_nonfinite(const double *x, long long n, int *flag) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
    if (!isfinite(x[t])) *flag = 1;
}

unsigned nb1(long long n) { return (unsigned)std::max(1LL, (n + 255) / 256); }

}  // namespace
}  // namespace fpc

using namespace fpc;

#define LAUNCH1(n) nb1(n), 256, 0, stream()
#define DONE()              \
  do {                      \
    FPC_LAUNCH_CHECK();     \
    FPC_HIP(hipStreamSynchronize(stream())); \
    return 0;               \
  } while (0)

extern "C" {

int fpc_apply_bc_cleanup(double *u, double *v, double *w, const unsigned char *au, const unsigned char *av,
                         const unsigned char *aw, const double *uin, int nx, int ny, int nz) {
  long long N = std::max((long long)(nx + 1) * ny * nz, std::max((long long)nx * (ny + 1) * nz, (long long)nx * ny * (nz + 1)));
  k_bc<<<LAUNCH1(N)>>>(u, v, w, au, av, aw, uin, nx, ny, nz);
  DONE();
}

int fpc_conv_u(double *conv, const double *u, const double *v, const double *w, const unsigned char *a, int nx, int ny,
               int nz, double dx, double dy, double dz) {
  k_conv<<<LAUNCH1((long long)(nx + 1) * ny * nz)>>>(0, conv, u, v, w, a, nx, ny, nz, dx, dy, dz);
  DONE();
}
int fpc_conv_v(double *conv, const double *u, const double *v, const double *w, const unsigned char *a, int nx, int ny,
               int nz, double dx, double dy, double dz) {
  k_conv<<<LAUNCH1((long long)nx * (ny + 1) * nz)>>>(1, conv, u, v, w, a, nx, ny, nz, dx, dy, dz);
  DONE();
}
int fpc_conv_w(double *conv, const double *u, const double *v, const double *w, const unsigned char *a, int nx, int ny,
               int nz, double dx, double dy, double dz) {
  k_conv<<<LAUNCH1((long long)nx * ny * (nz + 1))>>>(2, conv, u, v, w, a, nx, ny, nz, dx, dy, dz);
  DONE();
}

int fpc_build_u_full(double *rhsf, double *x0, const double *u, const double *conv, const double *gpx,
                     const unsigned char *a, const int *id, const double *bc, const double *sc, int nx, int ny, int nz,
                     double rho, double dt, int zero) {
  k_build_full<<<LAUNCH1((long long)(nx - 1) * ny * nz)>>>(0, rhsf, x0, u, conv, gpx, a, id, bc, sc, nx, ny, nz, rho, dt, zero);
  DONE();
}
int fpc_build_v_full(double *rhsf, double *x0, const double *v, const double *conv, const double *gpy,
                     const unsigned char *a, const int *id, const double *bc, const double *sc, int nx, int ny, int nz,
                     double rho, double dt, int zero) {
  k_build_full<<<LAUNCH1((long long)nx * (ny - 1) * nz)>>>(1, rhsf, x0, v, conv, gpy, a, id, bc, sc, nx, ny, nz, rho, dt, zero);
  DONE();
}
int fpc_build_w_full(double *rhsf, double *x0, const double *w, const double *conv, const double *gpz,
                     const unsigned char *a, const int *id, const double *bc, const double *sc, int nx, int ny, int nz,
                     double rho, double dt, int zero) {
  k_build_full<<<LAUNCH1((long long)nx * ny * (nz - 1))>>>(2, rhsf, x0, w, conv, gpz, a, id, bc, sc, nx, ny, nz, rho, dt, zero);
  DONE();
}

int fpc_scatter_u(double *o, const double *x, const double *base, int nx, int ny, int nz) {
  k_scatter<<<LAUNCH1((long long)(nx + 1) * ny * nz)>>>(0, o, x, base, nx, ny, nz);
  DONE();
}
int fpc_scatter_v(double *o, const double *x, const double *base, int nx, int ny, int nz) {
  k_scatter<<<LAUNCH1((long long)nx * (ny + 1) * nz)>>>(1, o, x, base, nx, ny, nz);
  DONE();
}
int fpc_scatter_w(double *o, const double *x, const double *base, int nx, int ny, int nz) {
  k_scatter<<<LAUNCH1((long long)nx * ny * (nz + 1))>>>(2, o, x, base, nx, ny, nz);
  DONE();
}

int fpc_spmv_csr_1based(int nrows, const int *rowptr, const int *colind, const double *vals, const double *x, double *y) {
  if (nrows <= 0) return 0;
  k_csr<<<LAUNCH1(nrows)>>>(nrows, rowptr, colind, vals, x, y);
  DONE();
}

int fpc_pack_pressure_active(double *vec, const double *p, const unsigned char *fluid, const int *id, int N) {
  k_pack_masked<<<LAUNCH1(N)>>>(vec, p, fluid, id, N);
  DONE();
}
int fpc_unpack_gp_u(const double *gp, double *gpx, const int *id, int nx, int ny, int nz) {
  k_unpack_gp<<<LAUNCH1((long long)(nx + 1) * ny * nz)>>>(gp, gpx, id, (long long)(nx + 1) * ny * nz, 0);
  DONE();
}
int fpc_unpack_gp_v(const double *gp, double *gpy, const int *id, int nx, int ny, int nz, int nQu) {
  k_unpack_gp<<<LAUNCH1((long long)nx * (ny + 1) * nz)>>>(gp, gpy, id, (long long)nx * (ny + 1) * nz, nQu);
  DONE();
}
int fpc_unpack_gp_w(const double *gp, double *gpz, const int *id, int nx, int ny, int nz, int nQu, int nQv) {
  k_unpack_gp<<<LAUNCH1((long long)nx * ny * (nz + 1))>>>(gp, gpz, id, (long long)nx * ny * (nz + 1), (long long)nQu + nQv);
  DONE();
}
int fpc_pack_projection_q_u(double *q, const double *u, const int *id, int nx, int ny, int nz) {
  k_pack_q<<<LAUNCH1((long long)(nx + 1) * ny * nz)>>>(q, u, id, (long long)(nx + 1) * ny * nz, 0);
  DONE();
}
int fpc_pack_projection_q_v(double *q, const double *v, const int *id, int nx, int ny, int nz, int nQu) {
  k_pack_q<<<LAUNCH1((long long)nx * (ny + 1) * nz)>>>(q, v, id, (long long)nx * (ny + 1) * nz, nQu);
  DONE();
}
int fpc_pack_projection_q_w(double *q, const double *w, const int *id, int nx, int ny, int nz, int nQu, int nQv) {
  k_pack_q<<<LAUNCH1((long long)nx * ny * (nz + 1))>>>(q, w, id, (long long)nx * ny * (nz + 1), (long long)nQu + nQv);
  DONE();
}
int fpc_build_pressure_rhs_full(double *rhsf, double *x0, const double *Dqt, const double *dbc,
                                const unsigned char *fluid, const int *id, int N, double rho, double dt, double ramp) {
  k_prhs<<<LAUNCH1(N)>>>(rhsf, x0, Dqt, dbc, fluid, id, N, rho, dt, ramp);
  DONE();
}
int fpc_pack_phi_active(double *phi, const double *xp, const unsigned char *fluid, const int *id, int N) {
  k_pack_masked<<<LAUNCH1(N)>>>(phi, xp, fluid, id, N);
  DONE();
}
int fpc_apply_q_correction(double *q, const double *g, double factor, int n1) {
  k_qcorr<<<LAUNCH1(n1)>>>(q, g, factor, n1);
  DONE();
}
int fpc_unpack_projection_q_u(const double *q, double *u, const int *id, int nx, int ny, int nz) {
  k_unpack_q<<<LAUNCH1((long long)(nx + 1) * ny * nz)>>>(q, u, id, (long long)(nx + 1) * ny * nz, 0);
  DONE();
}
int fpc_unpack_projection_q_v(const double *q, double *v, const int *id, int nx, int ny, int nz, int nQu) {
  k_unpack_q<<<LAUNCH1((long long)nx * (ny + 1) * nz)>>>(q, v, id, (long long)nx * (ny + 1) * nz, nQu);
  DONE();
}
int fpc_unpack_projection_q_w(const double *q, double *w, const int *id, int nx, int ny, int nz, int nQu, int nQv) {
  k_unpack_q<<<LAUNCH1((long long)nx * ny * (nz + 1))>>>(q, w, id, (long long)nx * ny * (nz + 1), (long long)nQu + nQv);
  DONE();
}
int fpc_add_pressure_phi(double *p, const double *phi, const unsigned char *fluid, const int *id, int N) {
  k_add_phi<<<LAUNCH1(N)>>>(p, phi, fluid, id, N);
  DONE();
}
int fpc_zero_doubles(double *p, long long n) {
  if (n <= 0) return 0;
  k_zero<<<(unsigned)std::min(65536LL, nb1(n) * 1LL), 256, 0, stream()>>>(p, n);
  DONE();
}


}  // extern "C"

// ---------------------------------------------------------------- fused stages (asynchronous variants are internal)
namespace fpc {
int stage_explicit_async(const fpc_flow_ops *o, double *u, double *v, double *w, const double *p,
                         const double *uin, double *ru, double *xu, double *rv, double *xv, double *rw, double *xw,
                         int zero) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  long long N = std::max((long long)(nx + 1) * ny * nz, std::max((long long)nx * (ny + 1) * nz, (long long)nx * ny * (nz + 1)));
  k_bc<<<LAUNCH1(N)>>>(u, v, w, o->active_u, o->active_v, o->active_w, uin, nx, ny, nz);
  FPC_LAUNCH_CHECK();
  double *outs[3][2] = {{ru, xu}, {rv, xv}, {rw, xw}};
  for (int c = 0; c < 3; ++c) {
    Dims d;
    d.nx = nx - (c == 0); d.ny = ny - (c == 1); d.nz = nz - (c == 2);
    dim3 g, b;
    launch_shape(d, g, b);
    {
      // bytes per box row: the three velocity fields and p once (shared reads cached),
      // masks (4 bytes), scale and bc (16), rhs and x0 (16)
      FPC_PROF_BEGIN("explicit");
      k_explicit<<<g, b, 0, stream()>>>(c, *o, u, v, w, p, outs[c][0], outs[c][1], zero);
      FPC_PROF_END("explicit", (double)d.nx * d.ny * d.nz * 68.0);
    }
    FPC_LAUNCH_CHECK();
  }
  return 0;
}

int stage_scatter_async(const fpc_flow_ops *o, const double *xu, const double *xv, const double *xw, const double *u,
                        const double *v, const double *w, double *ut, double *vt, double *wt, const double *uin) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  k_scatter<<<LAUNCH1((long long)(nx + 1) * ny * nz)>>>(0, ut, xu, u, nx, ny, nz);
  k_scatter<<<LAUNCH1((long long)nx * (ny + 1) * nz)>>>(1, vt, xv, v, nx, ny, nz);
  k_scatter<<<LAUNCH1((long long)nx * ny * (nz + 1))>>>(2, wt, xw, w, nx, ny, nz);
  FPC_LAUNCH_CHECK();
  long long N = std::max((long long)(nx + 1) * ny * nz, std::max((long long)nx * (ny + 1) * nz, (long long)nx * ny * (nz + 1)));
  k_bc<<<LAUNCH1(N)>>>(ut, vt, wt, o->active_u, o->active_v, o->active_w, uin, nx, ny, nz);
  FPC_LAUNCH_CHECK();
  return 0;
}

int stage_divergence_async(const fpc_flow_ops *o, const double *ut, const double *vt, const double *wt, double ramp,
                           double *rhs, double *x0) {
  Dims d;
  d.nx = o->nx; d.ny = o->ny; d.nz = o->nz;
  dim3 g, b;
  launch_shape(d, g, b);
  {
    // bytes per cell: three face velocities, fluid mask, rhs and x0
    FPC_PROF_BEGIN("divergence");
    k_divergence<<<g, b, 0, stream()>>>(*o, ut, vt, wt, ramp, rhs, x0);
    FPC_PROF_END("divergence", (double)d.nx * d.ny * d.nz * 41.0);
  }
  FPC_LAUNCH_CHECK();
  return 0;
}

int stage_projection_async(const fpc_flow_ops *o, const double *phi, const double *ut, const double *vt,
                           const double *wt, double *u, double *v, double *w, double *p, const double *uin,
                           double *qout) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  long long N = std::max((long long)(nx + 1) * ny * nz, std::max((long long)nx * (ny + 1) * nz, (long long)nx * ny * (nz + 1)));
  {
    // bytes per face (three components): u~ read, u written, phi read once per cell, masks
    FPC_PROF_BEGIN("projection");
    k_project<<<LAUNCH1(N)>>>(*o, phi, ut, vt, wt, u, v, w, qout);
    FPC_PROF_END("projection", (double)N * 3.0 * 18.0 + (double)nx * ny * nz * 9.0);
  }
  FPC_LAUNCH_CHECK();
  k_bc<<<LAUNCH1(N)>>>(u, v, w, o->active_u, o->active_v, o->active_w, uin, nx, ny, nz);
  FPC_LAUNCH_CHECK();
  long long Np = (long long)nx * ny * nz;
  k_add_phi_box<<<LAUNCH1(Np)>>>(p, phi, o->fluid_p, Np);
  FPC_LAUNCH_CHECK();
  return 0;
}

// The two diagnostics kernels with their partial sums in part (kDiagBlocks * 6 doubles).
static int diagnostics_launch(const fpc_flow_ops *o, const double *u, const double *v, const double *w,
                              const double *p, const double *qout, double ramp, int Np, double *out, double *part) {
  const int NBD = diag_blocks(o->nx, o->ny, o->nz);
  k_diag_partials<<<NBD, 256, 0, stream()>>>(*o, u, v, w, p, qout, ramp, part);
  FPC_LAUNCH_CHECK();
  k_diag_final<<<1, 1024, 0, stream()>>>(part, NBD, Np, out);
  FPC_LAUNCH_CHECK();
  return 0;
}

int stage_diagnostics_async(const fpc_flow_ops *o, const double *u, const double *v, const double *w, const double *p,
                            const double *qout, double ramp, int Np, double *out) {
  double *part = scratch_doubles((size_t)kDiagBlocks * 6, 6);
  if (!part) return FPC_ERROR_MEMORY;
  return diagnostics_launch(o, u, v, w, p, qout, ramp, Np, out, part);
}

namespace {
// Device address of `scratch` on each GPU, looked up once per device on its first call.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

int scratch_on_current_device(double **ptr) {
  int dev = -1;
  FPC_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error(FPC_ERROR_GENERIC, "device %d beyond the %d devices of the scratch table", dev, kMaxDevices);
    return FPC_ERROR_GENERIC;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *a = nullptr;
    hipError_t e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !a) {
      set_error(FPC_ERROR_GENERIC, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
      return FPC_ERROR_GENERIC;
    }
    g_scratch_ptr[dev] = (double *)a;
  }
  *ptr = g_scratch_ptr[dev];
  return 0;
}
}  // namespace
}  // namespace fpc

extern "C" {

int fpc_stage_explicit(const fpc_flow_ops *o, double *u, double *v, double *w, const double *p, const double *uin,
                       double *ru, double *xu, double *rv, double *xv, double *rw, double *xw, int zero) {
  int rc = stage_explicit_async(o, u, v, w, p, uin, ru, xu, rv, xv, rw, xw, zero);
  if (rc) return rc;
  FPC_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int fpc_stage_scatter(const fpc_flow_ops *o, const double *xu, const double *xv, const double *xw, const double *u,
                      const double *v, const double *w, double *ut, double *vt, double *wt, const double *uin) {
  int rc = stage_scatter_async(o, xu, xv, xw, u, v, w, ut, vt, wt, uin);
  if (rc) return rc;
  FPC_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int fpc_stage_divergence(const fpc_flow_ops *o, const double *ut, const double *vt, const double *wt, double ramp,
                         double *rhs, double *x0) {
  int rc = stage_divergence_async(o, ut, vt, wt, ramp, rhs, x0);
  if (rc) return rc;
  FPC_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int fpc_stage_projection(const fpc_flow_ops *o, const double *phi, const double *ut, const double *vt,
                         const double *wt, double *u, double *v, double *w, double *p, const double *uin,
                         double *qout) {
  int rc = stage_projection_async(o, phi, ut, vt, wt, u, v, w, p, uin, qout);
  if (rc) return rc;
  FPC_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int fpc_stage_diagnostics(const fpc_flow_ops *o, const double *u, const double *v, const double *w, const double *p,
                          const double *qout, double ramp, int Np, double *out) {
  return stage_diagnostics_async(o, u, v, w, p, qout, ramp, Np, out);
}

int fpc_stage_diagnostics_noalloc(const fpc_flow_ops *o, const double *u, const double *v, const double *w,
                                  const double *p, const double *qout, double ramp, int Np, double *out) {
  FPC_CHECK_ARG(o && u && v && w && p && qout && out, "null argument");
  FPC_CHECK_ARG(o->nx > 0 && o->ny > 0 && o->nz > 0 && Np > 0, "empty grid");
  // the stream must exist already: creating the library stream here would allocate
  FPC_CHECK_ARG(stream_ready(), "no stream yet: call fpc_Initialize or fpc_SetStream first");
  const int NBD = diag_blocks(o->nx, o->ny, o->nz);
  FPC_CHECK_ARG(NBD >= 1 && NBD <= kDiagBlocks, "diagnostics partition beyond the 2048-block scratch");
  double *part = nullptr;
  int rc = scratch_on_current_device(&part);
  if (rc) return rc;
  return diagnostics_launch(o, u, v, w, p, qout, ramp, Np, out, part);
}

int fpc_any_nonfinite(const double *x, long long n, int *flag) {
  FPC_CHECK_ARG(flag, "null flag");
  int *d = (int *)scratch_doubles(1, 7);
  if (!d) return FPC_ERROR_MEMORY;
  FPC_HIP(hipMemsetAsync(d, 0, sizeof(int), stream()));
  if (n > 0) {
    k_nonfinite<<<(unsigned)std::min(4096LL, nb1(n) * 1LL), 256, 0, stream()>>>(x, n, d);
    FPC_LAUNCH_CHECK();
  }
  FPC_HIP(hipMemcpyAsync(flag, d, sizeof(int), hipMemcpyDeviceToHost, stream()));
  FPC_HIP(hipStreamSynchronize(stream()));
  return 0;
}

}  // extern "C"

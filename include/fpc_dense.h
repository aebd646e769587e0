/*
 * fpc_dense.h -- C ABI of the per-timestep dense operations (libfpc.so).
 *
 * Two layers, both taking device pointers and launching on the library stream
 * (fpc_GetStream); each returns 0 or an FPC_ERROR_* code and has finished
 * when it returns unless stated otherwise.
 *
 *  1. One function per reference kernel of gpu_dense_ops.hpp, with the same
 *     argument meaning (full staggered arrays x fastest, 1-based active
 *     vectors with slot 0 unused), for callers that drive the reference's
 *     pack / CSR / unpack sequence (main.cu:256-350).
 *  2. Fused structured stages that compute the same values without the
 *     1-based active vectors or CSR: D and G are applied as MAC stencils in
 *     the CSR column order, so the sums round the same way.  The build's
 *     driver uses these.
 */
#ifndef FPC_DENSE_H
#define FPC_DENSE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- 1:1 counterparts of gpu_dense_ops.hpp ---- */
/* k_apply_bc_cleanup (gpu_dense_ops.hpp:5-10).  The outflow copy u[nx] = u[nx-1]
 * uses the cleaned last interior face (deterministic). */
int fpc_apply_bc_cleanup(double *u, double *v, double *w, const unsigned char *active_u,
                         const unsigned char *active_v, const unsigned char *active_w, const double *uin_step,
                         int nx, int ny, int nz);
/* k_conv_u / k_conv_v / k_conv_w (gpu_dense_ops.hpp:12-14) */
int fpc_conv_u(double *conv, const double *u, const double *v, const double *w, const unsigned char *active_u,
               int nx, int ny, int nz, double dx, double dy, double dz);
int fpc_conv_v(double *conv, const double *u, const double *v, const double *w, const unsigned char *active_v,
               int nx, int ny, int nz, double dx, double dy, double dz);
int fpc_conv_w(double *conv, const double *u, const double *v, const double *w, const unsigned char *active_w,
               int nx, int ny, int nz, double dx, double dy, double dz);
/* k_build_{u,v,w}_full (gpu_dense_ops.hpp:16-18) */
int fpc_build_u_full(double *rhsf, double *x0, const double *u, const double *conv, const double *gpx,
                     const unsigned char *active_u, const int *id_u, const double *rhs_bc, const double *scale_u,
                     int nx, int ny, int nz, double rho, double dt, int zero);
int fpc_build_v_full(double *rhsf, double *x0, const double *v, const double *conv, const double *gpy,
                     const unsigned char *active_v, const int *id_v, const double *rhs_bc, const double *scale_v,
                     int nx, int ny, int nz, double rho, double dt, int zero);
int fpc_build_w_full(double *rhsf, double *x0, const double *w, const double *conv, const double *gpz,
                     const unsigned char *active_w, const int *id_w, const double *rhs_bc, const double *scale_w,
                     int nx, int ny, int nz, double rho, double dt, int zero);
/* k_scatter_{u,v,w} (gpu_dense_ops.hpp:20-22); base may be NULL */
int fpc_scatter_u(double *uo, const double *x, const double *ubase, int nx, int ny, int nz);
int fpc_scatter_v(double *vo, const double *x, const double *vbase, int nx, int ny, int nz);
int fpc_scatter_w(double *wo, const double *x, const double *wbase, int nx, int ny, int nz);
/* k_spmv_csr_1based (gpu_dense_ops.hpp:61-68): y[row+1] = sum vals * x[colind] */
int fpc_spmv_csr_1based(int nrows, const int *rowptr, const int *colind, const double *vals, const double *x,
                        double *y);
/* pack / unpack / update kernels (gpu_dense_ops.hpp:70-106) */
int fpc_pack_pressure_active(double *vec, const double *p, const unsigned char *fluid_p, const int *id_p, int N);
int fpc_unpack_gp_u(const double *gp, double *gpx, const int *id_qu, int nx, int ny, int nz);
int fpc_unpack_gp_v(const double *gp, double *gpy, const int *id_qv, int nx, int ny, int nz, int nQu);
int fpc_unpack_gp_w(const double *gp, double *gpz, const int *id_qw, int nx, int ny, int nz, int nQu, int nQv);
int fpc_pack_projection_q_u(double *q, const double *u, const int *id_qu, int nx, int ny, int nz);
int fpc_pack_projection_q_v(double *q, const double *v, const int *id_qv, int nx, int ny, int nz, int nQu);
int fpc_pack_projection_q_w(double *q, const double *w, const int *id_qw, int nx, int ny, int nz, int nQu, int nQv);
int fpc_build_pressure_rhs_full(double *rhsf, double *x0, const double *Dqt, const double *div_bc_unit,
                                const unsigned char *fluid_p, const int *id_p, int N, double rho, double dt,
                                double ramp);
int fpc_pack_phi_active(double *phi, const double *xp_full, const unsigned char *fluid_p, const int *id_p, int N);
int fpc_apply_q_correction(double *q, const double *gphi, double factor, int nplus1);
int fpc_unpack_projection_q_u(const double *q, double *u, const int *id_qu, int nx, int ny, int nz);
int fpc_unpack_projection_q_v(const double *q, double *v, const int *id_qv, int nx, int ny, int nz, int nQu);
int fpc_unpack_projection_q_w(const double *q, double *w, const int *id_qw, int nx, int ny, int nz, int nQu,
                              int nQv);
int fpc_add_pressure_phi(double *p, const double *phi, const unsigned char *fluid_p, const int *id_p, int N);
/* cudaMemset(ptr, 0, n*8) of the driver (main.cu:262-266, 314, 318, 338, 340, 354) */
int fpc_zero_doubles(double *p, long long n);

/* ---- fused structured stages ---- */
/* Operator data of one channel, all device pointers (filled by fpc_sim's setup or
 * by the caller).  dcoef holds, per 2-D column (i + nx*j), the six divergence
 * coefficients [west, east, south, north, bottom, top] = alpha/(beta*h) with the
 * reference signs (immersed_operators.hpp:213-246), 0 where the face is not in D. */
typedef struct {
  int nx, ny, nz;
  double dx, dy, dz, rho, dt;
  const unsigned char *active_u, *active_v, *active_w, *fluid_p;
  const unsigned char *active_qu, *active_qv, *active_qw;
  const double *scale_u, *scale_v, *scale_w; /* per Struct-box row, 0 on inactive rows */
  const double *rhs_u_bc;                    /* per u box row (unit inlet amplitude, unramped) */
  const double *dcoef;                       /* 6 per (i, j) column */
  const double *div_bc;                      /* ny*nz plane: div_bc_unit of the i = 0 cells */
  int n_fu, n_fv, n_px, n_py;                /* force functional lists */
  const long long *fu_idx, *fv_idx, *px_idx, *py_idx;
  const double *fu_w, *fv_w, *px_w, *py_w;
} fpc_flow_ops;

/* BC cleanup, centred convection, G p on the faces and the momentum rhs / x0
 * boxes (main.cu:256-274). */
int fpc_stage_explicit(const fpc_flow_ops *ops, double *u, double *v, double *w, const double *p,
                       const double *uin_step, double *rhs_u, double *x_u, double *rhs_v, double *x_v,
                       double *rhs_w, double *x_w, int zero_guess);
/* u~ = u with the solved interior faces, then BC cleanup (main.cu:307-312). */
int fpc_stage_scatter(const fpc_flow_ops *ops, const double *x_u, const double *x_v, const double *x_w,
                      const double *u, const double *v, const double *w, double *ut, double *vt, double *wt,
                      const double *uin_step);
/* rhs_p = -(rho/dt)(D q~ + ramp div_bc) on fluid cells, 0 elsewhere; x_p = 0 (main.cu:314-321). */
int fpc_stage_divergence(const fpc_flow_ops *ops, const double *ut, const double *vt, const double *wt, double ramp,
                         double *rhs_p, double *x_p);
/* q = q~ - (dt/rho) G phi on the projection faces, u = u~ elsewhere, BC cleanup,
 * p += phi (main.cu:338-350).  qout (ny*nz) receives the projected outflow face
 * before the cleanup overwrites it (needed by the divergence diagnostic). */
int fpc_stage_projection(const fpc_flow_ops *ops, const double *phi, const double *ut, const double *vt,
                         const double *wt, double *u, double *v, double *w, double *p, const double *uin_step,
                         double *qout);
/* Diagnostics on the device (main.cu:354-363) and the drag/lift functional:
 * out[0] = divL2, out[1..3] = max|u|, max|v|, max|w|, out[4] = Fx, out[5] = Fy.
 * out is a device array of 8 doubles; asynchronous (no host wait). */
int fpc_stage_diagnostics(const fpc_flow_ops *ops, const double *u, const double *v, const double *w,
                          const double *p, const double *qout, double ramp, int Np, double *out);
/* fpc_stage_diagnostics without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  The block partial sums live in a static
 * device array of the library (2048 blocks x 6 doubles = 96 KiB per GPU, held until the
 * process ends); every grid fits, as the partition never exceeds 2048 blocks (a larger
 * one is rejected with FPC_ERROR_ARG).  The stream must exist already (fpc_Initialize
 * or fpc_SetStream), else FPC_ERROR_ARG.  All calls on one GPU share that array: calls
 * that overlap on one GPU (two streams, two host threads) race.  Same results, bit for
 * bit, as fpc_stage_diagnostics; asynchronous. */
int fpc_stage_diagnostics_noalloc(const fpc_flow_ops *ops, const double *u, const double *v, const double *w,
                                  const double *p, const double *qout, double ramp, int Np, double *out);
/* Non-finite check of a device vector (any_nonfinite_raw0, solver_helpers.hpp:331-335):
 * *flag (host) = 1 if any entry is NaN or Inf. */
int fpc_any_nonfinite(const double *x, long long n, int *flag);

#ifdef __cplusplus
}
#endif
#endif /* FPC_DENSE_H */

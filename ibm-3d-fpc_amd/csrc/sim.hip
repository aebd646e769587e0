// sim.hip -- time-step driver of the channel solver (libfpc).
//
// Mirrors main.cu:8-424 of the reference: one-time setup (fpc_setup), the four
// Struct grids, one 7-point stencil, four matrices set once from the embedded
// AoS boxes, GMRES + Jacobi for u, v, w and GMRES + PFMG for p with the
// reference's settings, then per step the phases of main.cu:249-376.
// Differences, all on purpose:
//  * the dense ops are the fused structured stages (no 1-based active vectors,
//    no CSR pack/unpack, no memsets); they write the solver vectors in place
//    instead of SetVector3D copies;
//  * diagnostics (divL2, max|u|, forces) are reduced on the device;
//  * the non-finite checks run on the device (exit codes 911 / 912 become
//    error returns);
//  * the pressure solve keeps its preconditioned vectors (flexible GMRES form),
//    which saves the V-cycle of HYPRE's final update x += M^-1 (V y);
//  * with mg_precision = 32 the pressure solve runs the FP32 PFMG, and the
//    momentum Jacobi preconditioners keep 1/diag in FP32 (no reference
//    counterpart).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "fpc_dense.h"
#include "fpc_internal.hpp"
#include "fpc_setup.hpp"
#include "fpc_sim.h"
#include "fpc_struct.h"
#include "batched.hpp"
#include "sim_internal.hpp"

namespace fpc {
namespace {

__global__ void k_scale_inlet(const double *uin, double *uin_step, double ramp, int n) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) uin_step[t] = ramp * uin[t];
}

template <class T>
int upload(T **d, const std::vector<T> &h) {
  FPC_HIP(hipMalloc((void **)d, std::max<size_t>(1, h.size()) * sizeof(T)));
  if (!h.empty()) {
    FPC_HIP(hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, fpc::stream()));
    FPC_HIP(hipStreamSynchronize(fpc::stream()));
  }
  return 0;
}

}  // namespace
}  // namespace fpc

using namespace fpc;

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static int setup_solver_vel(fpc_StructSolver *s, fpc_StructSolver *pc, const fpc_params &par, fpc_StructMatrix A,
                            fpc_StructVector b, fpc_StructVector x) {
  // main.cu:174-183
  int rc = fpc_StructGMRESCreate(s);
  rc |= fpc_StructGMRESSetTol(*s, par.vel_tol);
  rc |= fpc_StructGMRESSetMaxIter(*s, par.vel_maxit);
  rc |= fpc_StructGMRESSetKDim(*s, par.vel_kdim);
  rc |= fpc_StructGMRESSetPrintLevel(*s, par.vel_print);
  rc |= fpc_StructJacobiCreate(pc);
  rc |= fpc_StructJacobiSetMaxIter(*pc, par.vel_jacobi_iters);
  rc |= fpc_StructJacobiSetZeroGuess(*pc);
  // mixed precision: the momentum Jacobi preconditioner on its stored FP32 1/diag
  rc |= fpc_StructJacobiSetPrecision(*pc, par.mg_precision == 32 ? 32 : 64);
  rc |= fpc_StructGMRESSetPrecond(*s, fpc_StructJacobiSolve, fpc_StructJacobiSetup, *pc);
  if (rc) return rc;
  return fpc_StructGMRESSetup(*s, A, b, x);
}

static int setup_solver_p(fpc_StructSolver *s, fpc_StructSolver *pc, const fpc_params &par, fpc_StructMatrix A,
                          fpc_StructVector b, fpc_StructVector x) {
  // main.cu:207-221.  The preconditioned vectors of the Arnoldi steps are kept (flexible
  // form): the same iterates as HYPRE's right-preconditioned GMRES, whose final update
  // x += M^-1 (V y) costs one more V-cycle per solve; here x += Z y.  FP32 V-cycles
  // need the flexible form in any case.
  const bool flex = par.mg_precision == 32 || fpc::tile_tuning().p_flexible;
  int rc = flex ? fpc_StructFlexGMRESCreate(s) : fpc_StructGMRESCreate(s);
  rc |= fpc_StructGMRESSetTol(*s, par.p_tol);
  rc |= fpc_StructGMRESSetMaxIter(*s, par.p_maxit);
  rc |= fpc_StructGMRESSetKDim(*s, par.p_kdim);
  rc |= fpc_StructGMRESSetPrintLevel(*s, par.p_print);
  rc |= fpc_StructGMRESSetLogging(*s, 1);
  rc |= fpc_StructPFMGCreate(pc);
  rc |= fpc_StructPFMGSetMaxIter(*pc, par.p_pc_maxit);
  rc |= fpc_StructPFMGSetTol(*pc, 0.0);
  rc |= fpc_StructPFMGSetRelaxType(*pc, par.p_pc_relax_type);
  rc |= fpc_StructPFMGSetRAPType(*pc, par.p_pc_rap_type);
  rc |= fpc_StructPFMGSetPrintLevel(*pc, 0);
  rc |= fpc_StructPFMGSetZeroGuess(*pc);
  rc |= fpc_StructPFMGSetPrecision(*pc, par.mg_precision == 32 ? 32 : 64);
  rc |= fpc_StructGMRESSetPrecond(*s, fpc_StructPFMGSolve, fpc_StructPFMGSetup, *pc);
  if (rc) return rc;
  return fpc_StructGMRESSetup(*s, A, b, x);
}

int fpc::sim_run_phase(fpc_sim M, int ph) {
  const fpc_params &par = M->par;
  int nx = par.nx, ny = par.ny, nz = par.nz;
  hipStream_t st = fpc::stream();
  int rc = 0;
  switch (ph) {
    case 0: {  // main.cu:250-251
      M->step += 1;
      M->ramp = std::min(1.0, (double)M->step / (double)par.inlet_ramp_steps);
      int n = ny * nz;
      k_scale_inlet<<<(n + 255) / 256, 256, 0, st>>>(M->uin, M->uin_step, M->ramp, n);
      FPC_LAUNCH_CHECK();
      return 0;
    }
    case 1: {  // main.cu:253-281
      double t0 = now_s();
      rc = stage_explicit_async(&M->ops, M->u, M->v, M->w, M->p, M->uin_step, M->pbu, M->pxu, M->pbv, M->pxv, M->pbw,
                                M->pxw, par.use_zero_initial_guess != 0);
      if (rc) return rc;
      if (par.use_zero_initial_guess)  // x0 = 0 just written: the solves copy b as r0
        for (fpc_StructSolver s : {M->su, M->sv, M->sw}) fpc::gmres_next_guess_is_zero(s);
      FPC_HIP(hipStreamSynchronize(st));
      M->t_build += now_s() - t0;
      return 0;
    }
    case 2: {  // main.cu:283-305
      double t0 = now_s();
      fpc_StructSolver s3[3] = {M->su, M->sv, M->sw};
      fpc_StructMatrix A3[3] = {M->Au, M->Av, M->Aw};
      fpc_StructVector b3[3] = {M->bu, M->bv, M->bw}, x3[3] = {M->xu, M->xv, M->xw};
      // the three solves in lockstep: each computes what it computes alone, with the
      // host syncs of the three shared
      int rc3[3] = {0, 0, 0};
      rc = gmres_solve_many(s3, A3, b3, x3, 3, rc3);
      if (rc) return rc;
      for (int c = 0; c < 3; ++c) {
        M->err[c] = rc3[c];
        if (rc3[c] & FPC_ERROR_CONV) fpc_ClearError(FPC_ERROR_CONV);  // common.hpp:39
        fpc_StructGMRESGetNumIterations(s3[c], &M->its[c]);
        fpc_StructGMRESGetFinalRelativeResidualNorm(s3[c], &M->rel[c]);
      }
      M->t_mom += now_s() - t0;
      if (!M->nf_mom) {
        const std::vector<const double *> xs = {M->pxu, M->pxv, M->pxw};
        const std::vector<long long> nb = {(long long)(nx - 1) * ny * nz, (long long)nx * (ny - 1) * nz,
                                           (long long)nx * ny * (nz - 1)};
        rc = nonfinite_create(xs, nb, &M->nf_mom);
        if (rc) return rc;
      }
      rc = nonfinite_enqueue(M->nf_mom);
      if (rc) return rc;
      FPC_HIP(hipStreamSynchronize(st));
      const int *f = nonfinite_flags(M->nf_mom);
      if (f[0] | f[1] | f[2]) {
        fpc::set_error(FPC_ERROR_GENERIC, "non-finite momentum solve output at step %d (reference exit 911)", M->step);
        return 911;
      }
      return 0;
    }
    case 3:  // main.cu:307-312
      rc = stage_scatter_async(&M->ops, M->pxu, M->pxv, M->pxw, M->u, M->v, M->w, M->ut, M->vt, M->wt, M->uin_step);
      return rc;
    case 4:  // main.cu:314-324
      rc = stage_divergence_async(&M->ops, M->ut, M->vt, M->wt, M->ramp, M->pbp, M->pxp);
      if (rc) return rc;
      fpc::gmres_next_guess_is_zero(M->sp);  // x0 = 0 just written (main.cu:320)
      return 0;
    case 5: {  // main.cu:326-336
      double t0 = now_s();
      rc = fpc_StructGMRESSolve(M->sp, M->Ap, M->bp, M->xp);
      M->err[3] = rc;
      if (rc & ~FPC_ERROR_CONV) return rc;
      if (rc & FPC_ERROR_CONV) fpc_ClearError(FPC_ERROR_CONV);
      fpc_StructGMRESGetNumIterations(M->sp, &M->its[3]);
      fpc_StructGMRESGetFinalRelativeResidualNorm(M->sp, &M->rel[3]);
      M->t_p += now_s() - t0;
      int f = 0;
      rc = fpc_any_nonfinite(M->pxp, (long long)nx * ny * nz, &f);
      if (rc) return rc;
      if (f) {
        fpc::set_error(FPC_ERROR_GENERIC, "non-finite pressure solve output at step %d (reference exit 912)", M->step);
        return 912;
      }
      return 0;
    }
    case 6: {  // main.cu:338-351
      double t0 = now_s();
      rc = stage_projection_async(&M->ops, M->pxp, M->ut, M->vt, M->wt, M->u, M->v, M->w, M->p, M->uin_step, M->qout);
      if (rc) return rc;
      FPC_HIP(hipStreamSynchronize(st));
      M->t_proj += now_s() - t0;
      return 0;
    }
    case 7: {  // main.cu:352-368
      double t0 = now_s();
      rc = stage_diagnostics_async(&M->ops, M->u, M->v, M->w, M->p, M->qout, M->ramp, M->S.Np_, M->diag);
      if (rc) return rc;
      FPC_HIP(hipMemcpyAsync(M->hdiag, M->diag, 8 * sizeof(double), hipMemcpyDeviceToHost, st));
      FPC_HIP(hipStreamSynchronize(st));
      M->t_diag += now_s() - t0;
      for (int c = 0; c < 4; ++c) M->its_total[c] += M->its[c];
      return 0;
    }
    default:
      return FPC_ERROR_ARG;
  }
}

static double *field_ptr(fpc_sim M, const std::string &n, long long *len) {
  const fpc_params &P = M->par;
  long long Nu = (long long)(P.nx + 1) * P.ny * P.nz, Nv = (long long)P.nx * (P.ny + 1) * P.nz,
            Nw = (long long)P.nx * P.ny * (P.nz + 1), Np = (long long)P.nx * P.ny * P.nz;
  long long Bu = (long long)(P.nx - 1) * P.ny * P.nz, Bv = (long long)P.nx * (P.ny - 1) * P.nz,
            Bw = (long long)P.nx * P.ny * (P.nz - 1);
  struct E { const char *n; double *p; long long len; } tab[] = {
      {"u", M->u, Nu}, {"v", M->v, Nv}, {"w", M->w, Nw}, {"p", M->p, Np},
      {"u_tilde", M->ut, Nu}, {"v_tilde", M->vt, Nv}, {"w_tilde", M->wt, Nw},
      {"rhs_u", M->pbu, Bu}, {"x_u", M->pxu, Bu}, {"rhs_v", M->pbv, Bv}, {"x_v", M->pxv, Bv},
      {"rhs_w", M->pbw, Bw}, {"x_w", M->pxw, Bw}, {"rhs_p", M->pbp, Np}, {"x_p", M->pxp, Np},
      {"uin_step", M->uin_step, (long long)P.ny * P.nz}, {"qout", M->qout, (long long)P.ny * P.nz}};
  for (auto &e : tab)
    if (n == e.n) {
      *len = e.len;
      return e.p;
    }
  return nullptr;
}

extern "C" {

int fpc_sim_create(const fpc_params *par, fpc_sim *out) {
  FPC_CHECK_ARG(par && out, "null argument");
  FPC_CHECK_ARG(par->nx >= 2 && par->ny >= 2 && par->nz >= 2, "grid must be at least 2x2x2");
  ensure_init();
  fpc_sim M = new fpc_sim_s();
  M->par = *par;
  if (hipStreamCreate(&M->stream) != hipSuccess) {
    delete M;
    set_error(FPC_ERROR_GENERIC, "fpc_sim_create: hipStreamCreate failed");
    return FPC_ERROR_GENERIC;
  }
  fpc::StreamScope scope(M->stream);
  M->S.build(*par);
  const host::SetupData &S = M->S;
  int nx = par->nx, ny = par->ny, nz = par->nz, rc = 0;
  rc |= upload(&M->au, S.active_u);
  rc |= upload(&M->av, S.active_v);
  rc |= upload(&M->aw, S.active_w);
  rc |= upload(&M->fp, S.fluid_p);
  rc |= upload(&M->aqu, S.active_qu);
  rc |= upload(&M->aqv, S.active_qv);
  rc |= upload(&M->aqw, S.active_qw);
  rc |= upload(&M->scu, S.scale_u_box);
  rc |= upload(&M->scv, S.scale_v_box);
  rc |= upload(&M->scw, S.scale_w_box);
  rc |= upload(&M->bcu, S.rhs_u_bc_box);
  rc |= upload(&M->dcoef, S.dcoef);
  rc |= upload(&M->divbc, S.div_bc_plane);
  rc |= upload(&M->uin, S.uin);
  M->uin_base = S.uin;
  M->divbc_base = S.div_bc_plane;
  M->theta.assign((size_t)ny * nz, 1.0);
  rc |= upload(&M->uin_step, S.uin);
  rc |= upload(&M->fu_i, S.fu_idx);
  rc |= upload(&M->fv_i, S.fv_idx);
  rc |= upload(&M->px_i, S.px_idx);
  rc |= upload(&M->py_i, S.py_idx);
  rc |= upload(&M->fu_w, S.fu_w);
  rc |= upload(&M->fv_w, S.fv_w);
  rc |= upload(&M->px_w, S.px_w);
  rc |= upload(&M->py_w, S.py_w);
  std::vector<double> u0, v0, w0, p0;
  S.initial_fields(u0, v0, w0, p0);
  rc |= upload(&M->u, u0);
  rc |= upload(&M->v, v0);
  rc |= upload(&M->w, w0);
  rc |= upload(&M->p, p0);
  rc |= upload(&M->ut, u0);
  rc |= upload(&M->vt, v0);
  rc |= upload(&M->wt, w0);
  std::vector<double> zplane((size_t)ny * nz, 0.0), z8(8, 0.0);
  rc |= upload(&M->qout, zplane);
  rc |= upload(&M->diag, z8);
  if (rc) {
    fpc_sim_destroy(M);
    return rc;
  }
  if (hipHostMalloc((void **)&M->hdiag, 8 * sizeof(double)) != hipSuccess) {
    fpc_sim_destroy(M);
    return FPC_ERROR_MEMORY;
  }
  for (int q = 0; q < 8; ++q) M->hdiag[q] = 0.0;
  fpc_flow_ops &o = M->ops;
  o.nx = nx; o.ny = ny; o.nz = nz;
  o.dx = S.dx; o.dy = S.dy; o.dz = S.dz; o.rho = par->rho; o.dt = par->dt;
  o.active_u = M->au; o.active_v = M->av; o.active_w = M->aw; o.fluid_p = M->fp;
  o.active_qu = M->aqu; o.active_qv = M->aqv; o.active_qw = M->aqw;
  o.scale_u = M->scu; o.scale_v = M->scv; o.scale_w = M->scw;
  o.rhs_u_bc = M->bcu; o.dcoef = M->dcoef; o.div_bc = M->divbc;
  o.n_fu = (int)S.fu_idx.size(); o.n_fv = (int)S.fv_idx.size();
  o.n_px = (int)S.px_idx.size(); o.n_py = (int)S.py_idx.size();
  o.fu_idx = M->fu_i; o.fv_idx = M->fv_i; o.px_idx = M->px_i; o.py_idx = M->py_i;
  o.fu_w = M->fu_w; o.fv_w = M->fv_w; o.px_w = M->px_w; o.py_w = M->py_w;
  // Struct objects (main.cu:151-171)
  M->gu = fpc_CreateGrid3D(nx - 1, ny, nz);
  M->gv = fpc_CreateGrid3D(nx, ny - 1, nz);
  M->gw = fpc_CreateGrid3D(nx, ny, nz - 1);
  M->gp = fpc_CreateGrid3D(nx, ny, nz);
  M->st = fpc_CreateStencil3D();
  if (!M->gu || !M->gv || !M->gw || !M->gp || !M->st) {
    fpc_sim_destroy(M);
    return FPC_ERROR_GENERIC;
  }
  M->Au = fpc_CreateMatrix3D(M->gu, M->st, nx - 1, ny, nz, S.Au.data());
  M->Av = fpc_CreateMatrix3D(M->gv, M->st, nx, ny - 1, nz, S.Av.data());
  M->Aw = fpc_CreateMatrix3D(M->gw, M->st, nx, ny, nz - 1, S.Aw.data());
  M->Ap = fpc_CreateMatrix3D(M->gp, M->st, nx, ny, nz, S.Ap.data());
  M->bu = fpc_CreateVector3D(M->gu); M->xu = fpc_CreateVector3D(M->gu);
  M->bv = fpc_CreateVector3D(M->gv); M->xv = fpc_CreateVector3D(M->gv);
  M->bw = fpc_CreateVector3D(M->gw); M->xw = fpc_CreateVector3D(M->gw);
  M->bp = fpc_CreateVector3D(M->gp); M->xp = fpc_CreateVector3D(M->gp);
  if (!M->Au || !M->Av || !M->Aw || !M->Ap || !M->bu || !M->xu || !M->bv || !M->xv || !M->bw || !M->xw || !M->bp || !M->xp) {
    fpc_sim_destroy(M);
    return FPC_ERROR_GENERIC;
  }
  fpc_StructVectorGetDataPointer(M->bu, &M->pbu);
  fpc_StructVectorGetDataPointer(M->xu, &M->pxu);
  fpc_StructVectorGetDataPointer(M->bv, &M->pbv);
  fpc_StructVectorGetDataPointer(M->xv, &M->pxv);
  fpc_StructVectorGetDataPointer(M->bw, &M->pbw);
  fpc_StructVectorGetDataPointer(M->xw, &M->pxw);
  fpc_StructVectorGetDataPointer(M->bp, &M->pbp);
  fpc_StructVectorGetDataPointer(M->xp, &M->pxp);
  rc = setup_solver_vel(&M->su, &M->pu, *par, M->Au, M->bu, M->xu);
  if (!rc) rc = setup_solver_vel(&M->sv, &M->pv, *par, M->Av, M->bv, M->xv);
  if (!rc) rc = setup_solver_vel(&M->sw, &M->pw, *par, M->Aw, M->bw, M->xw);
  if (!rc) rc = setup_solver_p(&M->sp, &M->pp, *par, M->Ap, M->bp, M->xp);
  if (rc) {
    fpc_sim_destroy(M);
    return rc;
  }
  FPC_HIP(hipStreamSynchronize(fpc::stream()));
  *out = M;
  return 0;
}

int fpc_sim_phase(fpc_sim M, int phase) {
  FPC_CHECK_ARG(M, "null sim");
  fpc::StreamScope scope(M->stream);
  return fpc::sim_run_phase(M, phase);
}

int fpc_sim_step(fpc_sim M, int nsteps) {
  FPC_CHECK_ARG(M, "null sim");
  fpc::StreamScope scope(M->stream);
  for (int n = 0; n < nsteps; ++n)
    for (int ph = 0; ph < 8; ++ph) {
      int rc = fpc::sim_run_phase(M, ph);
      if (rc) return rc;
    }
  return 0;
}

long long fpc_sim_len(fpc_sim M, const char *name) {
  if (!M) return -1;
  long long n = -1;
  if (field_ptr(M, name, &n)) return n;
  const void *p;
  int es;
  if (M->S.view(name, &p, &n, &es) == 0) return n;
  return -1;
}

int fpc_sim_copy(fpc_sim M, const char *name, void *dst, long long nbytes) {
  FPC_CHECK_ARG(M && name && dst, "null argument");
  fpc::StreamScope scope(M->stream);
  long long n = -1;
  double *d = field_ptr(M, name, &n);
  if (d) {
    FPC_CHECK_ARG(nbytes >= n * (long long)sizeof(double), "destination too small");
    FPC_HIP(hipMemcpyAsync(dst, d, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, fpc::stream()));
    FPC_HIP(hipStreamSynchronize(fpc::stream()));
    return 0;
  }
  const void *p;
  int es;
  FPC_CHECK_ARG(M->S.view(name, &p, &n, &es) == 0, "unknown array name");
  FPC_CHECK_ARG(nbytes >= n * es, "destination too small");
  std::memcpy(dst, p, (size_t)(n * es));
  return 0;
}

int fpc_sim_set(fpc_sim M, const char *name, const void *src, long long nbytes) {
  FPC_CHECK_ARG(M && name && src, "null argument");
  fpc::StreamScope scope(M->stream);
  if (std::string(name) == "step") {
    FPC_CHECK_ARG(nbytes == (long long)sizeof(int), "step is one int");
    std::memcpy(&M->step, src, sizeof(int));
    return 0;
  }
  if (std::string(name) == "inlet_amplitude") {
    const size_t n = M->uin_base.size();
    FPC_CHECK_ARG(nbytes == (long long)(n * sizeof(double)), "size mismatch (ny*nz doubles)");
    std::memcpy(M->theta.data(), src, n * sizeof(double));
    std::vector<double> ui(n), db(n);
    for (size_t q = 0; q < n; ++q) {
      ui[q] = M->uin_base[q] * M->theta[q];
      db[q] = M->divbc_base[q] * M->theta[q];
    }
    FPC_HIP(hipMemcpyAsync(M->uin, ui.data(), n * sizeof(double), hipMemcpyHostToDevice, fpc::stream()));
    FPC_HIP(hipMemcpyAsync(M->divbc, db.data(), n * sizeof(double), hipMemcpyHostToDevice, fpc::stream()));
    FPC_HIP(hipStreamSynchronize(fpc::stream()));
    return 0;
  }
  long long n = -1;
  double *d = field_ptr(M, name, &n);
  FPC_CHECK_ARG(d, "unknown field");
  FPC_CHECK_ARG(nbytes == n * (long long)sizeof(double), "size mismatch");
  // a field set by the caller may be a solver's initial guess: no zero-guess shortcut
  for (fpc_StructSolver s : {M->su, M->sv, M->sw, M->sp}) fpc::gmres_clear_guess_hint(s);
  FPC_HIP(hipMemcpyAsync(d, src, (size_t)nbytes, hipMemcpyHostToDevice, fpc::stream()));
  FPC_HIP(hipStreamSynchronize(fpc::stream()));
  return 0;
}

int fpc_sim_device_ptr(fpc_sim M, const char *name, void **ptr) {
  FPC_CHECK_ARG(M && name && ptr, "null argument");
  long long n;
  double *d = field_ptr(M, name, &n);
  FPC_CHECK_ARG(d, "unknown field");
  *ptr = d;
  return 0;
}

void *fpc_sim_stream(fpc_sim M) { return M ? (void *)M->stream : nullptr; }

// The simulation's u, v, w, p on the host (after its stream's work).
static int fields_to_host(fpc_sim M, std::vector<double> *f) {
  const fpc_params &P = M->par;
  const size_t n[4] = {(size_t)(P.nx + 1) * P.ny * P.nz, (size_t)P.nx * (P.ny + 1) * P.nz,
                       (size_t)P.nx * P.ny * (P.nz + 1), (size_t)P.nx * P.ny * P.nz};
  const double *d[4] = {M->u, M->v, M->w, M->p};
  for (int c = 0; c < 4; ++c) {
    f[c].resize(n[c]);
    FPC_HIP(hipMemcpyAsync(f[c].data(), d[c], n[c] * sizeof(double), hipMemcpyDeviceToHost, M->stream));
  }
  FPC_HIP(hipStreamSynchronize(M->stream));
  return 0;
}

int fpc_sim_write_vtk(fpc_sim M, const char *filename) {
  FPC_CHECK_ARG(M && filename, "null argument");
  std::vector<double> f[4];
  int rc = fields_to_host(M, f);
  if (rc) return rc;
  const fpc_params &P = M->par;
  return fpc_write_vtk_cellcentered(filename, f[0].data(), f[1].data(), f[2].data(), f[3].data(), M->S.fluid_p.data(),
                                    P.nx, P.ny, P.nz, P.Lx, P.Ly, P.Lz);
}

int fpc_sim_write_fields(fpc_sim M, const char *path) {
  FPC_CHECK_ARG(M && path, "null argument");
  std::vector<double> f[4];
  int rc = fields_to_host(M, f);
  if (rc) return rc;
  const fpc_params &P = M->par;
  return fpc_write_fields(path, M->step, P.nx, P.ny, P.nz, f[0].data(), f[1].data(), f[2].data(), f[3].data());
}

int fpc_sim_flow_ops(fpc_sim M, fpc_flow_ops *ops) {
  FPC_CHECK_ARG(M && ops, "null argument");
  *ops = M->ops;
  return 0;
}

double fpc_sim_scalar(fpc_sim M, const char *name) {
  if (!M) return NAN;
  std::string n(name);
  const fpc_params &P = M->par;
  if (n == "step") return M->step;
  if (n == "ramp") return M->ramp;
  if (n == "div_l2") return M->hdiag[0];
  if (n == "umax") return M->hdiag[1];
  if (n == "vmax") return M->hdiag[2];
  if (n == "wmax") return M->hdiag[3];
  if (n == "fx") return M->hdiag[4];
  if (n == "fy") return M->hdiag[5];
  double ub = 4.0 * P.Um / 9.0, den = P.rho * ub * ub * (2.0 * P.Rcirc) * P.Lz;
  if (n == "cd") return 2.0 * M->hdiag[4] / den;
  if (n == "cl") return 2.0 * M->hdiag[5] / den;
  const char *counts[] = {"Np", "Nu", "Nv", "Nw", "Nqu", "Nqv", "Nqw", "n_pruned"};
  for (const char *c : counts)
    if (n == c) return (double)M->S.count(c);
  // hot copy of each operator (planned at its first solve): wall planes m (0: full
  // planes) and flagged points
  const char *zm[4] = {"zext_m_u", "zext_m_v", "zext_m_w", "zext_m_p"};
  const char *zf[4] = {"zext_flagged_u", "zext_flagged_v", "zext_flagged_w", "zext_flagged_p"};
  const fpc_StructMatrix ops[4] = {M->Au, M->Av, M->Aw, M->Ap};
  for (int c = 0; c < 4; ++c) {
    if (n != zm[c] && n != zf[c]) continue;
    int m = 0;
    long long f = 0;
    if (fpc_StructMatrixGetExtruded(ops[c], &m, &f)) return NAN;
    return n == zm[c] ? (double)m : (double)f;
  }
  const char *its[4] = {"its_u", "its_v", "its_w", "its_p"};
  const char *tot[4] = {"its_total_u", "its_total_v", "its_total_w", "its_total_p"};
  const char *rel[4] = {"relres_u", "relres_v", "relres_w", "relres_p"};
  const char *err[4] = {"err_u", "err_v", "err_w", "err_p"};
  for (int c = 0; c < 4; ++c) {
    if (n == its[c]) return M->its[c];
    if (n == tot[c]) return (double)M->its_total[c];
    if (n == rel[c]) return M->rel[c];
    if (n == err[c]) return M->err[c];
  }
  return NAN;
}

double fpc_sim_timer(fpc_sim M, const char *name) {
  if (!M) return NAN;
  std::string n(name);
  if (n == "t_build") return M->t_build;
  if (n == "t_mom") return M->t_mom;
  if (n == "t_p") return M->t_p;
  if (n == "t_proj") return M->t_proj;
  if (n == "t_diag") return M->t_diag;
  return NAN;
}

int fpc_sim_destroy(fpc_sim M) {
  if (!M) return 0;
  fpc::StreamScope scope(M->stream);
  (void)hipStreamSynchronize(M->stream);
  fpc_StructGMRESDestroy(M->su); fpc_StructJacobiDestroy(M->pu);
  fpc_StructGMRESDestroy(M->sv); fpc_StructJacobiDestroy(M->pv);
  fpc_StructGMRESDestroy(M->sw); fpc_StructJacobiDestroy(M->pw);
  fpc_StructGMRESDestroy(M->sp); fpc_StructPFMGDestroy(M->pp);
  for (fpc_StructVector x : {M->bu, M->xu, M->bv, M->xv, M->bw, M->xw, M->bp, M->xp}) fpc_StructVectorDestroy(x);
  for (fpc_StructMatrix A : {M->Au, M->Av, M->Aw, M->Ap}) fpc_StructMatrixDestroy(A);
  fpc_StructStencilDestroy(M->st);
  for (fpc_StructGrid g : {M->gu, M->gv, M->gw, M->gp}) fpc_StructGridDestroy(g);
  void *bufs[] = {M->au, M->av, M->aw, M->fp, M->aqu, M->aqv, M->aqw, M->scu, M->scv, M->scw, M->bcu, M->dcoef,
                  M->divbc, M->uin, M->uin_step, M->fu_i, M->fv_i, M->px_i, M->py_i, M->fu_w, M->fv_w, M->px_w,
                  M->py_w, M->u, M->v, M->w, M->p, M->ut, M->vt, M->wt, M->qout, M->diag};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (M->hdiag) (void)hipHostFree(M->hdiag);
  if (M->nf_mom) fpc::nonfinite_destroy(M->nf_mom);
  if (M->stream) (void)hipStreamDestroy(M->stream);
  delete M;
  return 0;
}

}  // extern "C"

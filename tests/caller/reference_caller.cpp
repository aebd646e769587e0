// reference_caller.cpp -- a caller written the way the reference's main() is
// (main.cu:8-404), run against libfpc: the HYPRE Struct names through
// fpc_hypre_compat.h for the grids, stencil, matrices, vectors and solvers
// (main.cu:151-221), and one fpc_dense.h layer-1 call per reference kernel
// launch, in the reference's order, for each time step (main.cu:249-363).
//
// What differs from main.cu, and why:
//  * the one-time CPU assembly comes from fpc_setup (libfpc's restatement of
//    immersed_geometry.hpp / immersed_operators.hpp / solver_helpers.hpp);
//  * MallocManaged* / CopyVecToManaged* return device memory, so every host
//    loop over those buffers in main.cu (initial copies :133, the inlet ramp
//    :251, the non-finite checks :302-305 and :333-336, the field read-back :352
//    and the diagnostics :358-363) becomes an explicit copy or fpc_any_nonfinite;
//  * <<<>>> launches become the fpc_* calls of the same name and arguments;
//    cudaMemset / cudaMemcpy become fpc_zero_doubles / fpc_MemcpyToDevice.
// Output: the reference's per-step line, then (argv[1]) a raw dump of u, v, w, p.
//
// Usage: reference_caller OUT.bin [reference flags]
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fpc_dense.h"
#include "fpc_hypre_compat.h"
#include "fpc_sim.h"

#define CALL(x)                                                                    \
  do {                                                                             \
    int rc_ = (x);                                                                 \
    if (rc_) {                                                                     \
      std::fprintf(stderr, "%s failed with %d: %s\n", #x, rc_, fpc_GetErrorMessage()); \
      return 1;                                                                    \
    }                                                                              \
  } while (0)

template <typename T>
static std::vector<T> setup_array(fpc_setup s, const char *name) {
  long long n = fpc_setup_len(s, name);
  std::vector<T> v(n > 0 ? (size_t)n : 0);
  if (n > 0) fpc_setup_copy(s, name, v.data(), n * (long long)sizeof(T));
  return v;
}

struct DeviceCSR {  // ManagedCSR (gpu_dense_ops.hpp:24-56) on device memory
  int nrows = 0;
  int *rowptr = nullptr, *colind = nullptr;
  double *vals = nullptr;
};

static DeviceCSR device_csr(fpc_setup s, const char *rp, const char *ci, const char *va) {
  DeviceCSR m;
  std::vector<int> r = setup_array<int>(s, rp), c = setup_array<int>(s, ci);
  std::vector<double> v = setup_array<double>(s, va);
  m.nrows = (int)r.size() - 1;
  CopyVecToManagedI(&m.rowptr, r);
  CopyVecToManagedI(&m.colind, c);
  CopyVecToManagedD(&m.vals, v);
  return m;
}

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s OUT.bin [flags]\n", argv[0]);
    return 2;
  }
  fpc_params par;
  fpc_default_params(&par);
  fpc_parse_args(argc - 1, argv + 1, &par);
  HYPRE_Initialize();
  HYPRE_SetMemoryLocation(HYPRE_MEMORY_DEVICE);
  HYPRE_SetExecutionPolicy(HYPRE_EXEC_DEVICE);
  const int nx = par.nx, ny = par.ny, nz = par.nz;
  const double dx = par.Lx / nx, dy = par.Ly / ny, dz = par.Lz / nz;

  fpc_setup S = nullptr;
  CALL(fpc_setup_create(&par, &S));
  const int Np = (int)fpc_setup_int(S, "Np"), Nqu = (int)fpc_setup_int(S, "Nqu"), Nqv = (int)fpc_setup_int(S, "Nqv"),
            Nqw = (int)fpc_setup_int(S, "Nqw");
  const int Nq = Nqu + Nqv + Nqw;
  auto u8 = [&](const char *n) {
    std::vector<unsigned char> a = setup_array<unsigned char>(S, n);
    return std::vector<char>(a.begin(), a.end());
  };
  std::vector<char> active_u = u8("active_u"), active_v = u8("active_v"), active_w = u8("active_w"),
                    fluid_p = u8("fluid_p");
  std::vector<int> id_u = setup_array<int>(S, "id_u"), id_v = setup_array<int>(S, "id_v"),
                   id_w = setup_array<int>(S, "id_w"), id_p = setup_array<int>(S, "id_p"),
                   id_qu = setup_array<int>(S, "id_qu"), id_qv = setup_array<int>(S, "id_qv"),
                   id_qw = setup_array<int>(S, "id_qw");
  std::vector<double> scale_u = setup_array<double>(S, "scale_u"), scale_v = setup_array<double>(S, "scale_v"),
                      scale_w = setup_array<double>(S, "scale_w"), rhs_u_bc_unit = setup_array<double>(S, "rhs_u_bc"),
                      rhs_v_bc = setup_array<double>(S, "rhs_v_bc"), rhs_w_bc = setup_array<double>(S, "rhs_w_bc"),
                      div_bc_unit = setup_array<double>(S, "div_bc"), uin = setup_array<double>(S, "uin");
  std::vector<double> Ahu_vals = setup_array<double>(S, "Au_vals"), Ahv_vals = setup_array<double>(S, "Av_vals"),
                      Ahw_vals = setup_array<double>(S, "Aw_vals"), Ahp_vals = setup_array<double>(S, "Ap_vals");

  HYPRE_Complex *Ahu_d = nullptr, *Ahv_d = nullptr, *Ahw_d = nullptr, *Ahp_d = nullptr;
  CopyVecToManagedC(&Ahu_d, Ahu_vals);
  CopyVecToManagedC(&Ahv_d, Ahv_vals);
  CopyVecToManagedC(&Ahw_d, Ahw_vals);
  CopyVecToManagedC(&Ahp_d, Ahp_vals);

  // main.cu:123-126: zero fields with the (unramped) inlet face, cleaned
  const size_t NuFull = (size_t)(nx + 1) * ny * nz, NvFull = (size_t)nx * (ny + 1) * nz,
               NwFull = (size_t)nx * ny * (nz + 1), NpFull = (size_t)nx * ny * nz;
  std::vector<double> u(NuFull, 0.0), v(NvFull, 0.0), w(NwFull, 0.0), p(NpFull, 0.0), uin_step = uin;
  double *du, *dv, *dw, *dp, *du_tilde, *dv_tilde, *dw_tilde, *duin_step, *dconv_u, *dconv_v, *dconv_w, *dgpx, *dgpy,
      *dgpz;
  CopyVecToManagedD(&du, u);
  CopyVecToManagedD(&dv, v);
  CopyVecToManagedD(&dw, w);
  CopyVecToManagedD(&dp, p);
  MallocManagedD(&du_tilde, NuFull);
  MallocManagedD(&dv_tilde, NvFull);
  MallocManagedD(&dw_tilde, NwFull);
  CopyVecToManagedD(&duin_step, uin_step);
  MallocManagedD(&dconv_u, NuFull);
  MallocManagedD(&dconv_v, NvFull);
  MallocManagedD(&dconv_w, NwFull);
  MallocManagedD(&dgpx, NuFull);
  MallocManagedD(&dgpy, NvFull);
  MallocManagedD(&dgpz, NwFull);
  unsigned char *dactive_u, *dactive_v, *dactive_w, *dfluid_p;
  int *did_u, *did_v, *did_w, *did_p, *did_qu, *did_qv, *did_qw;
  double *dscale_u, *dscale_v, *dscale_w, *drhs_u_bc_unit, *drhs_v_bc, *drhs_w_bc, *ddiv_bc_unit;
  CopyVecToManagedB(&dactive_u, active_u);
  CopyVecToManagedB(&dactive_v, active_v);
  CopyVecToManagedB(&dactive_w, active_w);
  CopyVecToManagedB(&dfluid_p, fluid_p);
  CopyVecToManagedI(&did_u, id_u);
  CopyVecToManagedI(&did_v, id_v);
  CopyVecToManagedI(&did_w, id_w);
  CopyVecToManagedI(&did_p, id_p);
  CopyVecToManagedI(&did_qu, id_qu);
  CopyVecToManagedI(&did_qv, id_qv);
  CopyVecToManagedI(&did_qw, id_qw);
  CopyVecToManagedD(&dscale_u, scale_u);
  CopyVecToManagedD(&dscale_v, scale_v);
  CopyVecToManagedD(&dscale_w, scale_w);
  CopyVecToManagedD(&drhs_u_bc_unit, rhs_u_bc_unit);
  CopyVecToManagedD(&drhs_v_bc, rhs_v_bc);
  CopyVecToManagedD(&drhs_w_bc, rhs_w_bc);
  CopyVecToManagedD(&ddiv_bc_unit, div_bc_unit);
  CALL(fpc_apply_bc_cleanup(du, dv, dw, dactive_u, dactive_v, dactive_w, duin_step, nx, ny, nz));
  DeviceCSR dDproj = device_csr(S, "D_rowptr", "D_colind", "D_vals");
  DeviceCSR dGproj = device_csr(S, "G_rowptr", "G_colind", "G_vals");

  HYPRE_Complex *rhsu_full, *xu_full, *rhsv_full, *xv_full, *rhsw_full, *xw_full, *rhsp_full, *xp_full;
  MallocManagedC(&rhsu_full, (size_t)(nx - 1) * ny * nz);
  MallocManagedC(&xu_full, (size_t)(nx - 1) * ny * nz);
  MallocManagedC(&rhsv_full, (size_t)nx * (ny - 1) * nz);
  MallocManagedC(&xv_full, (size_t)nx * (ny - 1) * nz);
  MallocManagedC(&rhsw_full, (size_t)nx * ny * (nz - 1));
  MallocManagedC(&xw_full, (size_t)nx * ny * (nz - 1));
  MallocManagedC(&rhsp_full, NpFull);
  MallocManagedC(&xp_full, NpFull);
  CALL(fpc_zero_doubles(rhsu_full, (long long)(nx - 1) * ny * nz));
  CALL(fpc_zero_doubles(xu_full, (long long)(nx - 1) * ny * nz));
  CALL(fpc_zero_doubles(rhsv_full, (long long)nx * (ny - 1) * nz));
  CALL(fpc_zero_doubles(xv_full, (long long)nx * (ny - 1) * nz));
  CALL(fpc_zero_doubles(rhsw_full, (long long)nx * ny * (nz - 1)));
  CALL(fpc_zero_doubles(xw_full, (long long)nx * ny * (nz - 1)));
  CALL(fpc_zero_doubles(rhsp_full, (long long)NpFull));
  CALL(fpc_zero_doubles(xp_full, (long long)NpFull));
  double *dpvec, *dgp, *dqtilde, *dDqt, *dphi_vec, *dgphi, *ddivv;
  MallocManagedD(&dpvec, (size_t)Np + 1);
  MallocManagedD(&dgp, (size_t)Nq + 1);
  MallocManagedD(&dqtilde, (size_t)Nq + 1);
  MallocManagedD(&dDqt, (size_t)Np + 1);
  MallocManagedD(&dphi_vec, (size_t)Np + 1);
  MallocManagedD(&dgphi, (size_t)Nq + 1);
  MallocManagedD(&ddivv, (size_t)Np + 1);

  // main.cu:151-221
  HYPRE_StructGrid gu = CreateGrid3D(nx - 1, ny, nz), gv = CreateGrid3D(nx, ny - 1, nz),
                   gw = CreateGrid3D(nx, ny, nz - 1), gpgrid = CreateGrid3D(nx, ny, nz);
  HYPRE_StructStencil st = CreateStencil3D();
  HYPRE_StructMatrix Ahu = CreateMatrix3D(gu, st, nx - 1, ny, nz, Ahu_d);
  HYPRE_StructMatrix Ahv = CreateMatrix3D(gv, st, nx, ny - 1, nz, Ahv_d);
  HYPRE_StructMatrix Ahw = CreateMatrix3D(gw, st, nx, ny, nz - 1, Ahw_d);
  HYPRE_StructMatrix Ahp = CreateMatrix3D(gpgrid, st, nx, ny, nz, Ahp_d);
  HYPRE_StructVector bhu = CreateVector3D(gu), xhu = CreateVector3D(gu);
  HYPRE_StructVector bhv = CreateVector3D(gv), xhv = CreateVector3D(gv);
  HYPRE_StructVector bhw = CreateVector3D(gw), xhw = CreateVector3D(gw);
  HYPRE_StructVector bhp = CreateVector3D(gpgrid), xhp = CreateVector3D(gpgrid);
  SetVector3D(bhu, nx - 1, ny, nz, rhsu_full);
  SetVector3D(xhu, nx - 1, ny, nz, xu_full);
  SetVector3D(bhv, nx, ny - 1, nz, rhsv_full);
  SetVector3D(xhv, nx, ny - 1, nz, xv_full);
  SetVector3D(bhw, nx, ny, nz - 1, rhsw_full);
  SetVector3D(xhw, nx, ny, nz - 1, xw_full);
  SetVector3D(bhp, nx, ny, nz, rhsp_full);
  SetVector3D(xhp, nx, ny, nz, xp_full);

  HYPRE_StructSolver solver[3], pcj[3], solver_p, pc_p;
  HYPRE_StructMatrix Am[3] = {Ahu, Ahv, Ahw};
  HYPRE_StructVector bm[3] = {bhu, bhv, bhw}, xm[3] = {xhu, xhv, xhw};
  for (int c = 0; c < 3; ++c) {
    CALL(HYPRE_StructGMRESCreate(MPI_COMM_WORLD, &solver[c]));
    CALL(HYPRE_StructGMRESSetTol(solver[c], par.vel_tol));
    CALL(HYPRE_StructGMRESSetMaxIter(solver[c], par.vel_maxit));
    CALL(HYPRE_StructGMRESSetKDim(solver[c], par.vel_kdim));
    CALL(HYPRE_StructGMRESSetPrintLevel(solver[c], par.vel_print));
    CALL(HYPRE_StructJacobiCreate(MPI_COMM_WORLD, &pcj[c]));
    CALL(HYPRE_StructJacobiSetMaxIter(pcj[c], par.vel_jacobi_iters));
    CALL(HYPRE_StructJacobiSetZeroGuess(pcj[c]));
    CALL(HYPRE_StructGMRESSetPrecond(solver[c], HYPRE_PtrToStructSolverFcn(HYPRE_StructJacobiSolve),
                                     HYPRE_PtrToStructSolverFcn(HYPRE_StructJacobiSetup), pcj[c]));
    CALL(HYPRE_StructGMRESSetup(solver[c], Am[c], bm[c], xm[c]));
  }
  CALL(HYPRE_StructGMRESCreate(MPI_COMM_WORLD, &solver_p));
  CALL(HYPRE_StructGMRESSetTol(solver_p, par.p_tol));
  CALL(HYPRE_StructGMRESSetMaxIter(solver_p, par.p_maxit));
  CALL(HYPRE_StructGMRESSetKDim(solver_p, par.p_kdim));
  CALL(HYPRE_StructGMRESSetPrintLevel(solver_p, par.p_print));
  CALL(HYPRE_StructGMRESSetLogging(solver_p, 1));
  CALL(HYPRE_StructPFMGCreate(MPI_COMM_WORLD, &pc_p));
  CALL(HYPRE_StructPFMGSetMaxIter(pc_p, par.p_pc_maxit));
  CALL(HYPRE_StructPFMGSetTol(pc_p, 0.0));
  CALL(HYPRE_StructPFMGSetRelaxType(pc_p, par.p_pc_relax_type));
  CALL(HYPRE_StructPFMGSetRAPType(pc_p, par.p_pc_rap_type));
  CALL(HYPRE_StructPFMGSetPrintLevel(pc_p, 0));
  CALL(HYPRE_StructPFMGSetZeroGuess(pc_p));
  CALL(HYPRE_StructGMRESSetPrecond(solver_p, HYPRE_PtrToStructSolverFcn(HYPRE_StructPFMGSolve),
                                   HYPRE_PtrToStructSolverFcn(HYPRE_StructPFMGSetup), pc_p));
  CALL(HYPRE_StructGMRESSetup(solver_p, Ahp, bhp, xhp));

  auto solve = [&](HYPRE_StructSolver s, HYPRE_StructMatrix A, HYPRE_StructVector b, HYPRE_StructVector x) {
    int rc = HYPRE_StructGMRESSolve(s, A, b, x);
    if (rc == HYPRE_ERROR_CONV) {  // HYPRE_SOLVE_CALL (common.hpp:36-43) clears it and goes on
      HYPRE_ClearError(HYPRE_ERROR_CONV);
      rc = 0;
    }
    return rc;
  };

  for (int step = 1; step <= par.nsteps; ++step) {
    const double ramp = std::min(1.0, (double)step / (double)par.inlet_ramp_steps);
    for (size_t t = 0; t < uin.size(); ++t) uin_step[t] = ramp * uin[t];
    CALL(fpc_MemcpyToDevice(duin_step, uin_step.data(), uin_step.size() * sizeof(double)));

    CALL(fpc_apply_bc_cleanup(du, dv, dw, dactive_u, dactive_v, dactive_w, duin_step, nx, ny, nz));
    CALL(fpc_conv_u(dconv_u, du, dv, dw, dactive_u, nx, ny, nz, dx, dy, dz));
    CALL(fpc_conv_v(dconv_v, du, dv, dw, dactive_v, nx, ny, nz, dx, dy, dz));
    CALL(fpc_conv_w(dconv_w, du, dv, dw, dactive_w, nx, ny, nz, dx, dy, dz));
    CALL(fpc_zero_doubles(dpvec, (long long)Np + 1));
    CALL(fpc_zero_doubles(dgp, (long long)Nq + 1));
    CALL(fpc_zero_doubles(dgpx, (long long)NuFull));
    CALL(fpc_zero_doubles(dgpy, (long long)NvFull));
    CALL(fpc_zero_doubles(dgpz, (long long)NwFull));
    CALL(fpc_pack_pressure_active(dpvec, dp, dfluid_p, did_p, nx * ny * nz));
    CALL(fpc_spmv_csr_1based(dGproj.nrows, dGproj.rowptr, dGproj.colind, dGproj.vals, dpvec, dgp));
    CALL(fpc_unpack_gp_u(dgp, dgpx, did_qu, nx, ny, nz));
    CALL(fpc_unpack_gp_v(dgp, dgpy, did_qv, nx, ny, nz, Nqu));
    CALL(fpc_unpack_gp_w(dgp, dgpz, did_qw, nx, ny, nz, Nqu, Nqv));
    const int zero = par.use_zero_initial_guess != 0;
    CALL(fpc_build_u_full(rhsu_full, xu_full, du, dconv_u, dgpx, dactive_u, did_u, drhs_u_bc_unit, dscale_u, nx, ny,
                          nz, par.rho, par.dt, zero));
    CALL(fpc_build_v_full(rhsv_full, xv_full, dv, dconv_v, dgpy, dactive_v, did_v, drhs_v_bc, dscale_v, nx, ny, nz,
                          par.rho, par.dt, zero));
    CALL(fpc_build_w_full(rhsw_full, xw_full, dw, dconv_w, dgpz, dactive_w, did_w, drhs_w_bc, dscale_w, nx, ny, nz,
                          par.rho, par.dt, zero));

    SetVector3D(bhu, nx - 1, ny, nz, rhsu_full);
    SetVector3D(xhu, nx - 1, ny, nz, xu_full);
    SetVector3D(bhv, nx, ny - 1, nz, rhsv_full);
    SetVector3D(xhv, nx, ny - 1, nz, xv_full);
    SetVector3D(bhw, nx, ny, nz - 1, rhsw_full);
    SetVector3D(xhw, nx, ny, nz - 1, xw_full);
    for (int c = 0; c < 3; ++c) CALL(solve(solver[c], Am[c], bm[c], xm[c]));
    int its[4] = {0, 0, 0, 0};
    double rel[4] = {0, 0, 0, 0};
    for (int c = 0; c < 3; ++c) {
      CALL(HYPRE_StructGMRESGetNumIterations(solver[c], &its[c]));
      CALL(HYPRE_StructGMRESGetFinalRelativeResidualNorm(solver[c], &rel[c]));
    }
    GetVector3D(xhu, nx - 1, ny, nz, xu_full);
    GetVector3D(xhv, nx, ny - 1, nz, xv_full);
    GetVector3D(xhw, nx, ny, nz - 1, xw_full);
    int bad[3] = {0, 0, 0};
    CALL(fpc_any_nonfinite(xu_full, (long long)(nx - 1) * ny * nz, &bad[0]));
    CALL(fpc_any_nonfinite(xv_full, (long long)nx * (ny - 1) * nz, &bad[1]));
    CALL(fpc_any_nonfinite(xw_full, (long long)nx * ny * (nz - 1), &bad[2]));
    if (bad[0] || bad[1] || bad[2]) {
      std::fprintf(stderr, "Non-finite momentum solve output at step %d\n", step);
      return 911;
    }

    CALL(fpc_scatter_u(du_tilde, xu_full, du, nx, ny, nz));
    CALL(fpc_scatter_v(dv_tilde, xv_full, dv, nx, ny, nz));
    CALL(fpc_scatter_w(dw_tilde, xw_full, dw, nx, ny, nz));
    CALL(fpc_apply_bc_cleanup(du_tilde, dv_tilde, dw_tilde, dactive_u, dactive_v, dactive_w, duin_step, nx, ny, nz));

    CALL(fpc_zero_doubles(dqtilde, (long long)Nq + 1));
    CALL(fpc_pack_projection_q_u(dqtilde, du_tilde, did_qu, nx, ny, nz));
    CALL(fpc_pack_projection_q_v(dqtilde, dv_tilde, did_qv, nx, ny, nz, Nqu));
    CALL(fpc_pack_projection_q_w(dqtilde, dw_tilde, did_qw, nx, ny, nz, Nqu, Nqv));
    CALL(fpc_zero_doubles(dDqt, (long long)Np + 1));
    CALL(fpc_spmv_csr_1based(dDproj.nrows, dDproj.rowptr, dDproj.colind, dDproj.vals, dqtilde, dDqt));
    CALL(fpc_build_pressure_rhs_full(rhsp_full, xp_full, dDqt, ddiv_bc_unit, dfluid_p, did_p, nx * ny * nz, par.rho,
                                     par.dt, ramp));
    SetVector3D(bhp, nx, ny, nz, rhsp_full);
    SetVector3D(xhp, nx, ny, nz, xp_full);
    CALL(solve(solver_p, Ahp, bhp, xhp));
    CALL(HYPRE_StructGMRESGetNumIterations(solver_p, &its[3]));
    CALL(HYPRE_StructGMRESGetFinalRelativeResidualNorm(solver_p, &rel[3]));
    GetVector3D(xhp, nx, ny, nz, xp_full);
    int badp = 0;
    CALL(fpc_any_nonfinite(xp_full, (long long)NpFull, &badp));
    if (badp) {
      std::fprintf(stderr, "Non-finite pressure solve output at step %d\n", step);
      return 912;
    }

    CALL(fpc_zero_doubles(dphi_vec, (long long)Np + 1));
    CALL(fpc_pack_phi_active(dphi_vec, xp_full, dfluid_p, did_p, nx * ny * nz));
    CALL(fpc_zero_doubles(dgphi, (long long)Nq + 1));
    CALL(fpc_spmv_csr_1based(dGproj.nrows, dGproj.rowptr, dGproj.colind, dGproj.vals, dphi_vec, dgphi));
    CALL(fpc_apply_q_correction(dqtilde, dgphi, par.dt / par.rho, Nq + 1));
    CALL(fpc_MemcpyToDevice(du, du_tilde, NuFull * sizeof(double)));
    CALL(fpc_MemcpyToDevice(dv, dv_tilde, NvFull * sizeof(double)));
    CALL(fpc_MemcpyToDevice(dw, dw_tilde, NwFull * sizeof(double)));
    CALL(fpc_unpack_projection_q_u(dqtilde, du, did_qu, nx, ny, nz));
    CALL(fpc_unpack_projection_q_v(dqtilde, dv, did_qv, nx, ny, nz, Nqu));
    CALL(fpc_unpack_projection_q_w(dqtilde, dw, did_qw, nx, ny, nz, Nqu, Nqv));
    CALL(fpc_apply_bc_cleanup(du, dv, dw, dactive_u, dactive_v, dactive_w, duin_step, nx, ny, nz));
    CALL(fpc_add_pressure_phi(dp, dphi_vec, dfluid_p, did_p, nx * ny * nz));
    CALL(fpc_MemcpyToHost(u.data(), du, NuFull * sizeof(double)));
    CALL(fpc_MemcpyToHost(v.data(), dv, NvFull * sizeof(double)));
    CALL(fpc_MemcpyToHost(w.data(), dw, NwFull * sizeof(double)));
    CALL(fpc_MemcpyToHost(p.data(), dp, NpFull * sizeof(double)));

    // main.cu:352-363: D q (projected) + ramp div_bc, then the maxima over the fields
    CALL(fpc_zero_doubles(ddivv, (long long)Np + 1));
    CALL(fpc_spmv_csr_1based(dDproj.nrows, dDproj.rowptr, dDproj.colind, dDproj.vals, dqtilde, ddivv));
    std::vector<double> divv((size_t)Np + 1);
    CALL(fpc_MemcpyToHost(divv.data(), ddivv, divv.size() * sizeof(double)));
    double div2 = 0.0;
    int ndiv = 0;
    for (int i = 1; i <= Np; ++i) {
      double d = divv[(size_t)i] + ramp * div_bc_unit[(size_t)i];
      div2 += d * d;
      ++ndiv;
    }
    const double div_l2 = std::sqrt(div2 / std::max(1, ndiv));
    double umax = 0.0, vmax = 0.0, wmax = 0.0;
    for (double x : u) umax = std::max(umax, std::abs(x));
    for (double x : v) vmax = std::max(vmax, std::abs(x));
    for (double x : w) wmax = std::max(wmax, std::abs(x));
    if (step == 1 || step == par.nsteps || (par.print_every > 0 && step % par.print_every == 0))
      std::printf("%4d  %.6f  %.6e  %.6e  %.6e  %.6e  (%d,%d,%d,%d)  (%.3e,%.3e,%.3e,%.3e)\n", step, ramp, umax, vmax,
                  wmax, div_l2, its[0], its[1], its[2], its[3], rel[0], rel[1], rel[2], rel[3]);
  }
  CALL(fpc_write_fields(argv[1], par.nsteps, nx, ny, nz, u.data(), v.data(), w.data(), p.data()));

  for (int c = 0; c < 3; ++c) {
    HYPRE_StructGMRESDestroy(solver[c]);
    HYPRE_StructJacobiDestroy(pcj[c]);
  }
  HYPRE_StructGMRESDestroy(solver_p);
  HYPRE_StructPFMGDestroy(pc_p);
  for (HYPRE_StructVector x : {bhu, xhu, bhv, xhv, bhw, xhw, bhp, xhp}) HYPRE_StructVectorDestroy(x);
  for (HYPRE_StructMatrix A : {Ahu, Ahv, Ahw, Ahp}) HYPRE_StructMatrixDestroy(A);
  HYPRE_StructStencilDestroy(st);
  for (HYPRE_StructGrid g : {gu, gv, gw, gpgrid}) HYPRE_StructGridDestroy(g);
  void *bufs[] = {rhsu_full, xu_full, rhsv_full, xv_full, rhsw_full, xw_full, rhsp_full, xp_full, Ahu_d, Ahv_d,
                  Ahw_d, Ahp_d, du, dv, dw, dp, du_tilde, dv_tilde, dw_tilde, duin_step, dconv_u, dconv_v, dconv_w,
                  dgpx, dgpy, dgpz, dactive_u, dactive_v, dactive_w, dfluid_p, did_u, did_v, did_w, did_p, did_qu,
                  did_qv, did_qw, dscale_u, dscale_v, dscale_w, drhs_u_bc_unit, drhs_v_bc, drhs_w_bc, ddiv_bc_unit,
                  dpvec, dgp, dqtilde, dDqt, dphi_vec, dgphi, ddivv, dDproj.rowptr, dDproj.colind, dDproj.vals,
                  dGproj.rowptr, dGproj.colind, dGproj.vals};
  for (void *b : bufs) fpc_Free(b);
  fpc_setup_destroy(S);
  HYPRE_Finalize();
  return 0;
}

// krylov.hip -- restarted (flexible) GMRES on the device and the Struct Jacobi
// preconditioner (libfpc).
//
// Replaces HYPRE_StructGMRES* (reference main.cu:174-221, 284-297, 327-331)
// and HYPRE_StructJacobi* (main.cu:179-182).  The control flow is HYPRE's
// published right-preconditioned GMRES: the same stopping test
// ||r|| <= max(atol, tol*||b||), the same Givens update, the same true-residual
// check and restart rules, and the same Givens-based residual at a full restart.
//
// MI355X mapping:
//  * all scalar state (Hessenberg matrix, rotations, residual estimate,
//    convergence flag) lives in device memory and is updated by one-block
//    kernels, so an Arnoldi step needs no host round trip;
//  * the host enqueues a batch of Arnoldi steps (sized by the previous solve's
//    count), then polls one small status record; kernels of steps past
//    convergence see the flag and return at once;
//  * classical Gram-Schmidt as one multi-dot pass plus one fused
//    update-and-norm pass (instead of 2j passes for modified Gram-Schmidt);
//  * basis vectors are stored unnormalised with a device scale factor, so no
//    separate scaling pass is needed;
//  * Jacobi (one sweep from zero: z = v / diag) is fused into the matvec;
//  * every reduction is two-stage over a partition fixed by the vector length
//    (red_blocks): run-to-run bitwise deterministic, and the same sums whether a
//    system is solved alone or in a batch (batched.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "fpc_internal.hpp"
#include "batched.hpp"
#include "krylov_device.hpp"
#include "stencil_tile.hpp"

namespace fpc {
namespace {

using namespace kry;

// w = A z with z = sigma * p (* dinv when Jacobi is fused), one thread per point.
// The current basis index is read from the device status.
template <typename DT>
__global__ void k_arnoldi_matvec(const double *__restrict__ c, long long pitch, int nx, int ny, int nz,
                                 const double *__restrict__ V, long long vp, const double *__restrict__ zin,
                                 const DT *__restrict__ dinv, State *S) {
  if (S->st.done) return;
  int cur = S->st.i;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  if (i >= nx || j >= ny) return;
  const double *z = zin ? zin : V + (long long)cur * vp;
  double sg = S->sigma[cur];
  double *w = (double *)V + (long long)(cur + 1) * vp;
  long long r = ((long long)k * ny + j) * nx + i;
  double s = arnoldi_row(c, pitch, FullMap(), nx, ny, nz, z, dinv, i, j, k, r);
  w[r] = sg * s;
}

template <bool kDinv, int TY, typename DT>
__global__ void __launch_bounds__(kTileX *TY)
    k_arnoldi_matvec_tiled(const double *__restrict__ c, long long pitch, ZMap zm, int nx, int ny, int nz,
                           const double *__restrict__ V, long long vp, const double *__restrict__ zin,
                           const DT *__restrict__ dinv, State *S, int kz) {
  if (S->st.done) return;
  const int cur = S->st.i;
  ArnoldiOp<kDinv, DT> op{c, pitch, nx, ny, zin ? zin : V + (long long)cur * vp, dinv,
                          (double *)V + (long long)(cur + 1) * vp, S->sigma[cur], zm};
  const int k0 = blockIdx.z * kz;
  march_tile<double, kTileX, TY>(op, nx, ny, nz, k0, min(k0 + kz, nz), blockIdx.x, blockIdx.y);
}

// w = sigma A z (z * 1/diag when jac: the FP64 or, DT = float, the FP32 1/diag)
// for the current basis index.  The z-marching kernels read the operator's hot
// copy when it has one; the one-point kernels (latency-bound boxes) the full planes.
template <typename DT>
int launch_arnoldi_matvec(const fpc_StructMatrix_s *A, const double *V, long long vp, const double *zin, bool jac,
                          State *S) {
  Dims d = dims_of(A->grid);
  if (tile_worthwhile(d)) {
    const Planes<double> P = planes64(A);
    const DT *dinv = jac ? jacobi_dinv<DT>(A) : nullptr;
    TileLaunch tl = tile_launch<8>(d);
    if (dinv)
      k_arnoldi_matvec_tiled<true, 8, DT><<<tl.grid, tl.block, 0, stream()>>>(P.c, P.pitch, P.zm, d.nx, d.ny, d.nz, V,
                                                                              vp, zin, dinv, S, tl.kz);
    else
      k_arnoldi_matvec_tiled<false, 8, DT><<<tl.grid, tl.block, 0, stream()>>>(P.c, P.pitch, P.zm, d.nx, d.ny, d.nz,
                                                                               V, vp, zin, nullptr, S, tl.kz);
  } else {
    const DT *dinv = jac ? jacobi_dinv<DT>(A, false) : nullptr;
    dim3 g, b;
    launch_shape(d, g, b);
    k_arnoldi_matvec<DT><<<g, b, 0, stream()>>>(A->coef, A->pitch, d.nx, d.ny, d.nz, V, vp, zin, dinv, S);
  }
  FPC_LAUNCH_CHECK();
  return 0;
}

// Partial dots <p_j, w> for j = 0..cur, eight vectors per pass over w.
__global__ void __launch_bounds__(256) k_multidot(const double *__restrict__ V, long long vp, long long n, State *S,
                                                  double *__restrict__ part) {
  if (S->st.done) return;
  multidot_body(V, vp, n, S, part, blockIdx.x, gridDim.x);
}

// h_{j,cur} = sigma_j * sum_b part[j][b]; one block per j.
__global__ void __launch_bounds__(256) k_reduce_h(State *S, const double *__restrict__ part, int nred) {
  reduce_h_body(S, part, nred, blockIdx.x);
}

// w -= sum_j h_j sigma_j p_j with partial sums of w^2.
__global__ void __launch_bounds__(256) k_orth_update(double *__restrict__ V, long long vp, long long n, State *S,
                                                     double *__restrict__ part) {
  if (S->st.done) return;
  orth_update_body(V, vp, n, S, part, blockIdx.x, gridDim.x);
}

// Norm of the new vector, Givens update, convergence flag (HYPRE's Arnoldi tail).
__global__ void __launch_bounds__(256) k_arnoldi_tail(State *S, const double *__restrict__ part, int nred) {
  arnoldi_tail_body(S, part, nred);
}

// Upper-triangular solve; y_j = sigma_j * rs_j (coefficients of the stored p_j).
__global__ void k_backsolve(State *S) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  backsolve_body(S);
}

// out (+)= [dinv *] sum_j y_j P_j   (x update or the preconditioner input)
template <typename DT>
__global__ void k_combine(const double *__restrict__ P, long long vp, long long n, const State *S,
                          const DT *__restrict__ dinv, double *__restrict__ out, int accumulate) {
  __shared__ double y[KM + 1];
  int i = S->st.i;
  for (int q = threadIdx.x; q < i; q += blockDim.x) y[q] = S->y[q];
  __syncthreads();
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    const double v = combine_point(P, vp, t, y, i, dinv);
    out[t] = accumulate ? out[t] + v : v;
  }
}

// HYPRE's residual at a full restart: rs_j rotated back, p_0 = sum_j rs_j v_j.
__global__ void k_restart_coefs(State *S) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  restart_coefs_body(S);
}

__global__ void k_restart_vec(double *__restrict__ V, long long vp, long long n, const State *S) {
  __shared__ double y[KM + 1];
  int i = S->st.i;
  for (int q = threadIdx.x; q <= i; q += blockDim.x) y[q] = S->y[q];
  __syncthreads();
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
    V[t] = restart_point(V, vp, t, y, i);
}

// Start of a cycle: i = 0, rs_0 = r_norm, sigma_0 = 1/r_norm.
__global__ void k_cycle_start(State *S, double r_norm, double eps, int iter, int maxit, int kdim) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  cycle_start_body(S, r_norm, eps, iter, maxit, kdim);
}

__global__ void k_scale_copy(double *__restrict__ dst, const double *__restrict__ src, const State *S, long long n) {
  int cur = S->st.i;
  double sg = S->sigma[cur];
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
    dst[t] = sg * src[t];
}

__global__ void k_set_done(State *S, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) S->st.done = v;
}

}  // namespace

// ---------------------------------------------------------------- Jacobi
namespace {
template <typename DT>
__global__ void k_jacobi_zero(const DT *__restrict__ dinv, const double *__restrict__ b, double *__restrict__ x,
                              long long n, const int *skip) {
  if (skip && *skip) return;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
    x[t] = b[t] * (double)dinv[t];
}
template <typename DT>
__global__ void k_jacobi_sweep(const double *__restrict__ c, long long pitch, int nx, int ny, int nz,
                               const DT *__restrict__ dinv, const double *__restrict__ b,
                               const double *__restrict__ x, double *__restrict__ xn, const int *skip) {
  if (skip && *skip) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  if (i >= nx || j >= ny) return;
  long long r = ((long long)k * ny + j) * nx + i, pl = (long long)nx * ny;
  double s = c[r] * x[r];
  if (i > 0) s += c[pitch + r] * x[r - 1];
  if (i < nx - 1) s += c[2 * pitch + r] * x[r + 1];
  if (j > 0) s += c[3 * pitch + r] * x[r - nx];
  if (j < ny - 1) s += c[4 * pitch + r] * x[r + nx];
  if (k > 0) s += c[5 * pitch + r] * x[r - pl];
  if (k < nz - 1) s += c[6 * pitch + r] * x[r + pl];
  xn[r] = x[r] + (b[r] - s) * (double)dinv[r];
}
unsigned nblk(long long n) {
  long long b = (n + 255) / 256;
  return (unsigned)std::max(1LL, std::min(b, 65536LL));
}
}  // namespace

int jacobi_apply(JacobiImpl *J, const double *b, double *x, const int *skip) {
  fpc_StructMatrix_s *A = J->A;
  int rc = ensure_dinv(A);
  if (rc) return rc;
  long long n = A->grid.N;
  if (J->bits == 32) {
    if ((rc = ensure_fp32(A))) return rc;
    k_jacobi_zero<float><<<nblk(n), 256, 0, stream()>>>(A->dinv32, b, x, n, skip);
  } else {
    k_jacobi_zero<double><<<nblk(n), 256, 0, stream()>>>(A->dinv, b, x, n, skip);
  }
  FPC_LAUNCH_CHECK();
  if (J->maxit > 1) {
    double *t = scratch_doubles((size_t)n, 3);
    if (!t) return FPC_ERROR_MEMORY;
    Dims d = dims_of(A->grid);
    dim3 g, bl;
    launch_shape(d, g, bl);
    for (int it = 1; it < J->maxit; ++it) {
      if (J->bits == 32)
        k_jacobi_sweep<float><<<g, bl, 0, stream()>>>(A->coef, A->pitch, d.nx, d.ny, d.nz, A->dinv32, b, x, t, skip);
      else
        k_jacobi_sweep<double><<<g, bl, 0, stream()>>>(A->coef, A->pitch, d.nx, d.ny, d.nz, A->dinv, b, x, t, skip);
      FPC_LAUNCH_CHECK();
      rc = vec_copy(x, t, n);
      if (rc) return rc;
    }
  }
  return 0;
}

// ---------------------------------------------------------------- GMRES host driver
namespace {

enum PcKind { PC_NONE = 0, PC_JACOBI1 = 1, PC_JACOBI = 2, PC_PFMG = 3, PC_CALLBACK = 4 };

struct GmresImpl {
  bool flexible = false;
  double tol = 1.0e-6, atol = 0.0;  // HYPRE defaults
  int maxit = 1000, kdim = 5, print = 0, logging = 0;
  fpc_PtrToStructSolverFcn pc_solve = nullptr, pc_setup = nullptr;
  fpc_StructSolver pc = nullptr;
  int pc_kind = PC_NONE;
  // workspace
  long long N = 0, vp = 0;
  int kd = 0;
  int nred = 1;  // blocks of the reductions over N points (red_blocks)
  double *V = nullptr;    // kd+1 basis vectors
  double *Z = nullptr;    // flexible: kd preconditioned vectors; standard: one work vector
  double *W = nullptr;    // work vector (true residual / precond input)
  double *part = nullptr; // NB * (KM + 1)
  double *dscal = nullptr;  // reduction outputs
  State *S = nullptr;
  Status *hst = nullptr;    // pinned
  double *hscal = nullptr;  // pinned
  int its = 0;
  double relres = 0.0;
  int pred = 4;  // expected Arnoldi steps per cycle (from the previous solve)
  bool x0_zero = false;  // the next solve starts from x = 0 (gmres_next_guess_is_zero)
  fpc_StructMatrix_s *A = nullptr;
  fpc_StructVector_s tmp_in, tmp_out;  // callback preconditioner views
};

void free_ws(GmresImpl *G) {
  if (G->V) (void)hipFree(G->V);
  if (G->Z) (void)hipFree(G->Z);
  if (G->W) (void)hipFree(G->W);
  if (G->part) (void)hipFree(G->part);
  if (G->dscal) (void)hipFree(G->dscal);
  if (G->S) (void)hipFree(G->S);
  if (G->hst) (void)hipHostFree(G->hst);
  if (G->hscal) (void)hipHostFree(G->hscal);
  G->V = G->Z = G->W = G->part = G->dscal = nullptr;
  G->S = nullptr;
  G->hst = nullptr;
  G->hscal = nullptr;
  G->N = 0;
}

GmresImpl *gimpl(fpc_StructSolver s) {
  if (!s || (s->kind != SK_GMRES && s->kind != SK_FGMRES)) return nullptr;
  return (GmresImpl *)s->impl;
}

int precond_apply(GmresImpl *G, const double *in, double *out, const int *skip) {
  switch (G->pc_kind) {
    case PC_JACOBI:
      return jacobi_apply((JacobiImpl *)G->pc->impl, in, out, skip);
    case PC_PFMG:
      return pfmg_apply(G->pc->impl, in, out, skip);
    case PC_CALLBACK: {
      G->tmp_in.data = (double *)in;
      G->tmp_out.data = out;
      FPC_HIP(hipStreamSynchronize(stream()));
      int rc = G->pc_solve(G->pc, G->A, &G->tmp_in, &G->tmp_out);
      return rc & ~FPC_ERROR_CONV;
    }
    default:
      return vec_copy(out, in, G->N);
  }
}

}  // namespace

void gmres_next_guess_is_zero(fpc_StructSolver s) {
  if (GmresImpl *G = gimpl(s)) G->x0_zero = true;
}

void gmres_clear_guess_hint(fpc_StructSolver s) {
  if (GmresImpl *G = gimpl(s)) G->x0_zero = false;
}

int gmres_batch_view(fpc_StructSolver s, GmresBatchView *v) {
  GmresImpl *G = gimpl(s);
  if (!G || !G->A || G->N == 0 || !G->S) {
    set_error(FPC_ERROR_ARG, "batched solve: a solver is not set up");
    return FPC_ERROR_ARG;
  }
  if (G->flexible && G->pc_kind != PC_PFMG) {
    set_error(FPC_ERROR_ARG, "batched solve: flexible GMRES is supported with the PFMG-style preconditioner only");
    return FPC_ERROR_ARG;
  }
  if (G->pc_kind == PC_NONE) v->pc = 0;
  else if (G->pc_kind == PC_JACOBI1 && ((JacobiImpl *)G->pc->impl)->bits == 64) v->pc = 1;
  else if (G->pc_kind == PC_PFMG) v->pc = 3;
  else {
    set_error(FPC_ERROR_ARG, "batched solve: only Jacobi (one sweep) and PFMG preconditioners are supported");
    return FPC_ERROR_ARG;
  }
  v->flexible = G->flexible ? 1 : 0;
  v->V = G->V;
  v->Z = G->Z;
  v->W = G->W;
  v->part = G->part;
  v->S = G->S;
  v->N = G->N;
  v->vp = G->vp;
  v->kd = G->kd;
  v->nred = G->nred;
  v->maxit = G->maxit;
  v->tol = G->tol;
  v->atol = G->atol;
  v->dinv = (v->pc == 1) ? G->A->dinv : nullptr;
  v->A = G->A;
  v->pfmg = (v->pc == 3) ? G->pc->impl : nullptr;
  v->pred = &G->pred;
  v->x0_zero = &G->x0_zero;
  v->its = &G->its;
  v->relres = &G->relres;
  return 0;
}

}  // namespace fpc

using namespace fpc;

extern "C" {

static int gmres_create(fpc_StructSolver *solver, bool flexible) {
  FPC_CHECK_ARG(solver, "null handle pointer");
  ensure_init();
  fpc_StructSolver s = new fpc_StructSolver_s();
  s->kind = flexible ? SK_FGMRES : SK_GMRES;
  GmresImpl *G = new GmresImpl();
  G->flexible = flexible;
  s->impl = G;
  *solver = s;
  return 0;
}

int fpc_StructGMRESCreate(fpc_StructSolver *solver) { return gmres_create(solver, false); }
int fpc_StructFlexGMRESCreate(fpc_StructSolver *solver) { return gmres_create(solver, true); }

#define GIMPL(s)                                   \
  GmresImpl *G = gimpl(s);                         \
  FPC_CHECK_ARG(G, "not a GMRES solver handle")

int fpc_StructGMRESSetTol(fpc_StructSolver s, double tol) { GIMPL(s); G->tol = tol; return 0; }
int fpc_StructGMRESSetAbsoluteTol(fpc_StructSolver s, double atol) { GIMPL(s); G->atol = atol; return 0; }
int fpc_StructGMRESSetMaxIter(fpc_StructSolver s, int maxit) { GIMPL(s); G->maxit = maxit; return 0; }
int fpc_StructGMRESSetKDim(fpc_StructSolver s, int kdim) {
  GIMPL(s);
  FPC_CHECK_ARG(kdim >= 1 && kdim <= kMaxKDim, "kdim must be 1..128");
  G->kdim = kdim;
  return 0;
}
int fpc_StructGMRESSetPrintLevel(fpc_StructSolver s, int level) { GIMPL(s); G->print = level; return 0; }
int fpc_StructGMRESSetLogging(fpc_StructSolver s, int logging) { GIMPL(s); G->logging = logging; return 0; }

int fpc_StructGMRESSetPrecond(fpc_StructSolver s, fpc_PtrToStructSolverFcn solve, fpc_PtrToStructSolverFcn setup,
                              fpc_StructSolver pc) {
  GIMPL(s);
  G->pc_solve = solve;
  G->pc_setup = setup;
  G->pc = pc;
  if (!solve) G->pc_kind = PC_NONE;
  else if (solve == fpc_StructJacobiSolve && pc && pc->kind == SK_JACOBI) G->pc_kind = PC_JACOBI;
  else if (solve == fpc_StructPFMGSolve && pc && pc->kind == SK_PFMG) G->pc_kind = PC_PFMG;
  else G->pc_kind = PC_CALLBACK;
  return 0;
}

int fpc_StructGMRESSetup(fpc_StructSolver s, fpc_StructMatrix A, fpc_StructVector b, fpc_StructVector x) {
  GIMPL(s);
  FPC_CHECK_ARG(A && b && x, "null argument");
  FPC_CHECK_ARG(A->initialized && b->initialized && x->initialized, "objects not initialized");
  FPC_CHECK_ARG(b->grid.N == A->grid.N && x->grid.N == A->grid.N, "size mismatch");
  G->A = A;
  int rc = ensure_dinv(A);
  if (rc) return rc;
  long long N = A->grid.N;
  if (G->N != N || G->kd != G->kdim) {
    free_ws(G);
    G->N = N;
    G->kd = G->kdim;
    G->vp = pad256(std::max(1LL, N), (int)sizeof(double));
    G->nred = red_blocks(N);
    FPC_HIP(hipMalloc(&G->V, (size_t)(G->kd + 1) * G->vp * sizeof(double)));
    int nz = G->flexible ? G->kd : 1;
    FPC_HIP(hipMalloc(&G->Z, (size_t)nz * G->vp * sizeof(double)));
    FPC_HIP(hipMalloc(&G->W, (size_t)G->vp * sizeof(double)));
    FPC_HIP(hipMalloc(&G->part, (size_t)NB * (KM + 1) * sizeof(double)));
    FPC_HIP(hipMalloc(&G->dscal, 8 * sizeof(double)));
    FPC_HIP(hipMalloc(&G->S, sizeof(State)));
    FPC_HIP(hipHostMalloc(&G->hst, sizeof(Status)));
    FPC_HIP(hipHostMalloc(&G->hscal, 8 * sizeof(double)));
  }
  G->tmp_in.grid = A->grid;
  G->tmp_out.grid = A->grid;
  G->tmp_in.initialized = G->tmp_out.initialized = true;
  // Jacobi with one sweep from zero is fused into the Arnoldi matvec.
  if (G->pc_kind == PC_JACOBI || G->pc_kind == PC_JACOBI1) {
    JacobiImpl *J = (JacobiImpl *)G->pc->impl;
    J->A = A;
    G->pc_kind = (J->maxit <= 1 && !G->flexible) ? PC_JACOBI1 : PC_JACOBI;
    if (J->bits == 32 && (rc = ensure_fp32(A))) return rc;
  }
  if (G->pc_kind == PC_PFMG) {
    rc = pfmg_setup(G->pc->impl, A);
    if (rc) return rc;
  }
  if (G->pc_kind == PC_CALLBACK && G->pc_setup) {
    rc = G->pc_setup(G->pc, A, b, x);
    if (rc & ~FPC_ERROR_CONV) return rc;
  }
  FPC_HIP(hipStreamSynchronize(stream()));
  return 0;
}

// ||b - A x|| into V[0] (or W), and optionally ||b||; the values land in hscal[0], hscal[1]
// once the stream has been synchronised (enqueued only).
static int true_residual(GmresImpl *G, const double *b, const double *x, double *rout, bool with_b) {
  int rc = residual_norm_partials(G->A, b, x, rout, G->part);
  if (rc) return rc;
  rc = reduce_partials(G->part, G->nred, G->dscal);
  if (rc) return rc;
  if (with_b) {
    rc = dot_partials(b, b, G->N, G->part + NB);
    if (!rc) rc = reduce_partials(G->part + NB, G->nred, G->dscal + 1);
    if (rc) return rc;
  }
  FPC_HIP(hipMemcpyAsync(G->hscal, G->dscal, 2 * sizeof(double), hipMemcpyDeviceToHost, stream()));
  return 0;
}

// One Arnoldi matvec with no preconditioner or with the fused one-sweep Jacobi.
static int arnoldi_step(GmresImpl *G) {
  const bool jac = G->pc_kind == PC_JACOBI1;
  const bool j32 = jac && ((JacobiImpl *)G->pc->impl)->bits == 32;
  // bytes per row: 7 coefficients, the input, 1/diag when Jacobi is fused (4 bytes in FP32), the output
  FPC_PROF_BEGIN(jac ? "spmv_jacobi" : "spmv");
  int rc = j32 ? launch_arnoldi_matvec<float>(G->A, G->V, G->vp, nullptr, true, G->S)
               : launch_arnoldi_matvec<double>(G->A, G->V, G->vp, nullptr, jac, G->S);
  FPC_PROF_END(jac ? "spmv_jacobi" : "spmv", stencil_bytes(G->A, jac ? (j32 ? 4 : 8) : 0, 16.0, true));
  return rc;
}

// step: host-side index of the current basis vector (for the byte counts only).
static int arnoldi_tail(GmresImpl *G, int step) {
  double nv = (double)(step + 1);
  {
    // bytes per row: the step+1 basis vectors once, w once per group of eight
    FPC_PROF_BEGIN("multidot");
    k_multidot<<<G->nred, 256, 0, stream()>>>(G->V, G->vp, G->N, G->S, G->part);
    FPC_PROF_END("multidot", (double)G->N * 8.0 * (nv + std::ceil(nv / 8.0)));
  }
  FPC_LAUNCH_CHECK();
  k_reduce_h<<<KM + 1, 256, 0, stream()>>>(G->S, G->part, G->nred);
  FPC_LAUNCH_CHECK();
  {
    // bytes per row: the step+1 basis vectors, w read and written
    FPC_PROF_BEGIN("orth_update");
    k_orth_update<<<G->nred, 256, 0, stream()>>>(G->V, G->vp, G->N, G->S, G->part);
    FPC_PROF_END("orth_update", (double)G->N * 8.0 * (nv + 2.0));
  }
  FPC_LAUNCH_CHECK();
  k_arnoldi_tail<<<1, 256, 0, stream()>>>(G->S, G->part, G->nred);
  FPC_LAUNCH_CHECK();
  return 0;
}

// One Arnoldi step when the preconditioner runs separately; step = host-side index
// of the current basis vector (used for the flexible Z slot).
static int arnoldi_step_pc(GmresImpl *G, int step) {
  int *skip = &G->S->st.done;
  double *z = G->flexible ? G->Z + (long long)step * G->vp : G->Z;
  // the preconditioner input is the stored (unnormalised) p_step
  int rc = precond_apply(G, G->V + (long long)step * G->vp, z, skip);
  if (rc) return rc;
  FPC_PROF_BEGIN("spmv");
  rc = launch_arnoldi_matvec<double>(G->A, G->V, G->vp, z, false, G->S);
  FPC_PROF_END("spmv", stencil_bytes(G->A, 0, 16.0, true));
  return rc;
}

// One Arnoldi step (preconditioner, matvec, orthogonalisation, Givens update).
static int run_step(GmresImpl *G, int step) {
  int rc = (G->pc_kind == PC_NONE || G->pc_kind == PC_JACOBI1) ? arnoldi_step(G) : arnoldi_step_pc(G, step);
  if (rc) return rc;
  return arnoldi_tail(G, step);
}

// One GMRES solve as a resumable run.  Each stage enqueues its launches and ends with
// the host copy its decision needs; after the stream is synchronised, run_advance reads
// that copy and enqueues the next stage.  The launches and the decisions are those of
// the solve loop in order (restarted GMRES with HYPRE's convergence test: true residual
// once the recursive one is below tol, stop when it is below tol or no longer falls),
// so runs advanced together -- one sync for all of them -- give the bits of the same
// runs made one after another.
struct GmresRun {
  enum Stage { kInit, kPre, kCycle, kPost, kDone };
  GmresImpl *G = nullptr;
  const double *b = nullptr;
  double *x = nullptr;
  Stage stage = kInit;
  double r_norm = 0.0, b_norm = 0.0, eps = 0.0, real_old = 0.0;
  int iter = 0, issued = 0, batch = 1;
  bool first_cycle = true;
  size_t prof_from = 0;
};

static unsigned combine_blocks(long long N) { return (unsigned)std::max(1LL, std::min((N + 255) / 256, 65536LL)); }

// A batch of Arnoldi steps and the copy of the status record.
static int run_batch(GmresRun &R) {
  GmresImpl *G = R.G;
  int nb = std::min(R.batch, G->kd - R.issued);
  nb = std::min(nb, G->maxit - R.iter - R.issued);
  if (nb <= 0) nb = 1;
  for (int q = 0; q < nb; ++q) {
    prof_tag(R.issued + q, G);
    int rc = run_step(G, R.issued + q);
    if (rc) return rc;
  }
  prof_tag(-1, nullptr);
  R.issued += nb;
  FPC_HIP(hipMemcpyAsync(G->hst, &G->S->st, sizeof(Status), hipMemcpyDeviceToHost, stream()));
  R.stage = GmresRun::kCycle;
  return 0;
}

// A new cycle: i = 0 from r_norm, then the first batch of steps (as many as the last
// solve's first cycle took).
static int run_cycle_start(GmresRun &R) {
  GmresImpl *G = R.G;
  k_cycle_start<<<1, 1, 0, stream()>>>(G->S, R.r_norm, R.eps, R.iter, G->maxit, G->kd);
  FPC_LAUNCH_CHECK();
  R.issued = 0;
  R.prof_from = prof_mark();
  R.batch = std::max(1, std::min(G->pred, G->kd));
  if (G->pc_kind == PC_CALLBACK) R.batch = 1;
  return run_batch(R);
}

// The top of the restart loop (host decisions only).
static int run_top(GmresRun &R) {
  GmresImpl *G = R.G;
  if (R.iter >= G->maxit || R.r_norm == 0.0) {
    R.stage = GmresRun::kDone;
    return 0;
  }
  if (R.r_norm <= R.eps) {  // check the true residual before another cycle
    R.stage = GmresRun::kPre;
    return true_residual(G, R.b, R.x, G->W, false);
  }
  return run_cycle_start(R);
}

static int run_begin(GmresRun &R) {
  R.stage = GmresRun::kInit;
  GmresImpl *G = R.G;
  if (G->x0_zero) {
    // x = 0: the initial residual is b (into V[0]), and ||r||^2 = ||b||^2 from the same partials
    G->x0_zero = false;
    int rc = copy_norm_partials(R.b, G->V, G->N, G->part);
    if (!rc) rc = reduce_partials(G->part, G->nred, G->dscal);
    if (!rc) rc = reduce_partials(G->part, G->nred, G->dscal + 1);
    if (rc) return rc;
    FPC_HIP(hipMemcpyAsync(G->hscal, G->dscal, 2 * sizeof(double), hipMemcpyDeviceToHost, stream()));
    return 0;
  }
  return true_residual(G, R.b, R.x, G->V, true);  // initial residual into V[0]
}

// After a sync: read this run's host copy and enqueue its next stage.
static int run_advance(GmresRun &R) {
  GmresImpl *G = R.G;
  const long long N = G->N;
  int rc = 0;
  switch (R.stage) {
    case GmresRun::kInit: {
      R.r_norm = std::sqrt(G->hscal[0]);
      R.b_norm = std::sqrt(G->hscal[1]);
      G->its = 0;
      if (!std::isfinite(R.b_norm) || !std::isfinite(R.r_norm)) {
        G->relres = NAN;
        R.stage = GmresRun::kDone;
        set_error(FPC_ERROR_GENERIC, "GMRES: non-finite right-hand side or initial residual");
        return FPC_ERROR_GENERIC;
      }
      const double den = (R.b_norm > 0.0) ? R.b_norm : R.r_norm;
      R.eps = std::max(G->atol, G->tol * den);
      R.real_old = R.b_norm;
      R.iter = 0;
      R.first_cycle = true;
      return run_top(R);
    }
    case GmresRun::kPre:
      R.r_norm = std::sqrt(G->hscal[0]);
      if (R.r_norm <= R.eps) {
        R.stage = GmresRun::kDone;
        return 0;
      }
      rc = vec_copy(G->V, G->W, N);
      if (rc) return rc;
      return run_cycle_start(R);
    case GmresRun::kCycle: {
      if (!G->hst->done && R.issued < G->kd) {
        R.batch = (G->pc_kind == PC_PFMG || G->pc_kind == PC_CALLBACK) ? 1 : 2;
        return run_batch(R);
      }
      const int i = G->hst->i;
      prof_drop_skipped(R.prof_from, i, G);
      R.iter = G->hst->iter;
      R.r_norm = G->hst->r_norm;
      if (R.first_cycle) G->pred = std::max(1, i);
      R.first_cycle = false;
      // x update
      fpc_StructMatrix_s *A = G->A;
      const unsigned nblk_n = combine_blocks(N);
      k_backsolve<<<1, 1, 0, stream()>>>(G->S);
      FPC_LAUNCH_CHECK();
      if (G->flexible) {
        k_combine<double><<<nblk_n, 256, 0, stream()>>>(G->Z, G->vp, N, G->S, nullptr, R.x, 1);
        FPC_LAUNCH_CHECK();
      } else if (G->pc_kind == PC_JACOBI1 && ((JacobiImpl *)G->pc->impl)->bits == 32) {
        k_combine<float><<<nblk_n, 256, 0, stream()>>>(G->V, G->vp, N, G->S, A->dinv32, R.x, 1);
        FPC_LAUNCH_CHECK();
      } else if (G->pc_kind == PC_JACOBI1 || G->pc_kind == PC_NONE) {
        k_combine<double><<<nblk_n, 256, 0, stream()>>>(G->V, G->vp, N, G->S,
                                                        G->pc_kind == PC_JACOBI1 ? A->dinv : nullptr, R.x, 1);
        FPC_LAUNCH_CHECK();
      } else {
        k_combine<double><<<nblk_n, 256, 0, stream()>>>(G->V, G->vp, N, G->S, nullptr, G->W, 0);
        FPC_LAUNCH_CHECK();
        rc = precond_apply(G, G->W, G->Z, nullptr);
        if (rc) return rc;
        rc = axpy_inplace(R.x, G->Z, N);
        if (rc) return rc;
      }
      if (R.r_norm <= R.eps) {
        R.stage = GmresRun::kPost;
        return true_residual(G, R.b, R.x, G->V, false);
      }
      // full cycle without convergence: Givens-based residual as the new p_0
      k_restart_coefs<<<1, 1, 0, stream()>>>(G->S);
      FPC_LAUNCH_CHECK();
      k_restart_vec<<<nblk_n, 256, 0, stream()>>>(G->V, G->vp, N, G->S);
      FPC_LAUNCH_CHECK();
      return run_top(R);
    }
    case GmresRun::kPost: {
      const double real_new = std::sqrt(G->hscal[0]);
      R.r_norm = real_new;
      if (R.r_norm <= R.eps || real_new >= R.real_old) {
        R.stage = GmresRun::kDone;
        return 0;
      }
      R.real_old = real_new;
      return run_top(R);  // V[0] holds the true residual
    }
    case GmresRun::kDone:
      return 0;
  }
  return 0;
}

// Checks and set-up of one solve (a PFMG hierarchy is rebuilt only when A changed).
static int run_prepare(GmresImpl *G, fpc_StructMatrix A, fpc_StructVector bv, fpc_StructVector xv) {
  FPC_CHECK_ARG(A && bv && xv && G->A == A && G->N == A->grid.N, "call Setup with this matrix first");
  FPC_CHECK_ARG(bv->grid.N == G->N && xv->grid.N == G->N, "size mismatch");
  int rc = ensure_dinv(A);
  if (rc) return rc;
  if (G->pc_kind == PC_PFMG) rc = pfmg_setup(G->pc->impl, A);  // no-op unless A changed
  return rc;
}

// Result of a finished run: iterations, relative residual, FPC_ERROR_CONV at maxit.
static int run_result(GmresRun &R) {
  GmresImpl *G = R.G;
  G->its = R.iter;
  G->relres = (R.b_norm > 0.0) ? R.r_norm / R.b_norm : R.r_norm;
  if (G->print > 0) std::printf("GMRES: %d iterations, relative residual %.6e\n", R.iter, G->relres);
  if (R.iter >= G->maxit && R.r_norm > R.eps && R.eps > 0) {
    set_error(FPC_ERROR_CONV, "GMRES: maximum iterations reached");
    return FPC_ERROR_CONV;
  }
  return 0;
}

}  // extern "C"

namespace fpc {

// n solves in lockstep on the calling thread's stream, one sync per round for all of
// them; rc[q] as fpc_StructGMRESSolve returns it for solve q.  A solve that fails
// outside convergence ends the call: its code is returned, the others are abandoned.
int gmres_solve_many(const fpc_StructSolver *s, const fpc_StructMatrix *A, const fpc_StructVector *b,
                     const fpc_StructVector *x, int n, int *rc) {
  std::vector<GmresRun> runs((size_t)n);
  for (int q = 0; q < n; ++q) {
    GmresRun &R = runs[(size_t)q];
    R.G = gimpl(s[q]);
    FPC_CHECK_ARG(R.G, "not a GMRES solver handle");
    int e = run_prepare(R.G, A[q], b[q], x[q]);
    if (e) return e;
    R.b = b[q]->data;
    R.x = x[q]->data;
  }
  for (GmresRun &R : runs) {
    int e = run_begin(R);
    if (e) return e;
  }
  for (;;) {
    bool live = false;
    for (const GmresRun &R : runs) live = live || R.stage != GmresRun::kDone;
    FPC_HIP(hipStreamSynchronize(stream()));
    if (!live) break;
    for (GmresRun &R : runs) {
      if (R.stage == GmresRun::kDone) continue;
      int e = run_advance(R);
      if (e) return e;
    }
  }
  for (int q = 0; q < n; ++q) rc[q] = run_result(runs[(size_t)q]);
  return 0;
}

}  // namespace fpc

extern "C" {

int fpc_StructGMRESSolve(fpc_StructSolver s, fpc_StructMatrix A, fpc_StructVector bv, fpc_StructVector xv) {
  GIMPL(s);
  (void)G;
  int r = 0;
  int rc = gmres_solve_many(&s, &A, &bv, &xv, 1, &r);
  return rc ? rc : r;
}

int fpc_StructGMRESGetNumIterations(fpc_StructSolver s, int *n) {
  GIMPL(s);
  FPC_CHECK_ARG(n, "null pointer");
  *n = G->its;
  return 0;
}

int fpc_StructGMRESGetFinalRelativeResidualNorm(fpc_StructSolver s, double *r) {
  GIMPL(s);
  FPC_CHECK_ARG(r, "null pointer");
  *r = G->relres;
  return 0;
}

int fpc_StructGMRESDestroy(fpc_StructSolver s) {
  if (!s) return 0;
  GmresImpl *G = gimpl(s);
  if (G) {
    free_ws(G);
    delete G;
  }
  delete s;
  return 0;
}

// ---------------------------------------------------------------- Jacobi handles
int fpc_StructJacobiCreate(fpc_StructSolver *solver) {
  FPC_CHECK_ARG(solver, "null handle pointer");
  fpc_StructSolver s = new fpc_StructSolver_s();
  s->kind = SK_JACOBI;
  s->impl = new JacobiImpl();
  *solver = s;
  return 0;
}

#define JIMPL(s)                                                              \
  FPC_CHECK_ARG(s && s->kind == SK_JACOBI, "not a Jacobi solver handle");     \
  JacobiImpl *J = (JacobiImpl *)s->impl

int fpc_StructJacobiSetMaxIter(fpc_StructSolver s, int maxit) { JIMPL(s); J->maxit = std::max(1, maxit); return 0; }
int fpc_StructJacobiSetTol(fpc_StructSolver s, double tol) { JIMPL(s); J->tol = tol; return 0; }
int fpc_StructJacobiSetZeroGuess(fpc_StructSolver s) { JIMPL(s); J->zero_guess = true; return 0; }
int fpc_StructJacobiSetPrecision(fpc_StructSolver s, int bits) {
  JIMPL(s);
  FPC_CHECK_ARG(bits == 32 || bits == 64, "bits must be 32 or 64");
  J->bits = bits;
  return 0;
}
int fpc_StructJacobiSetNonZeroGuess(fpc_StructSolver s) { JIMPL(s); J->zero_guess = false; return 0; }

int fpc_StructJacobiSetup(fpc_StructSolver s, fpc_StructMatrix A, fpc_StructVector b, fpc_StructVector x) {
  JIMPL(s);
  FPC_CHECK_ARG(A && A->initialized, "matrix not initialized");
  (void)b;
  (void)x;
  J->A = A;
  return ensure_dinv(A);
}

// Stand-alone Jacobi: maxit sweeps, from zero when ZeroGuess is set, else from x.
int fpc_StructJacobiSolve(fpc_StructSolver s, fpc_StructMatrix A, fpc_StructVector b, fpc_StructVector x) {
  JIMPL(s);
  FPC_CHECK_ARG(A && b && x && A->grid.N == b->grid.N && A->grid.N == x->grid.N, "bad arguments");
  J->A = A;
  int rc = ensure_dinv(A);
  if (rc) return rc;
  if (J->zero_guess) {
    rc = jacobi_apply(J, b->data, x->data, nullptr);
  } else {
    if (J->bits == 32 && (rc = ensure_fp32(A))) return rc;
    long long n = A->grid.N;
    double *t = scratch_doubles((size_t)n, 4);
    if (!t) return FPC_ERROR_MEMORY;
    Dims d = dims_of(A->grid);
    dim3 g, bl;
    launch_shape(d, g, bl);
    for (int it = 0; it < J->maxit; ++it) {
      if (J->bits == 32)
        k_jacobi_sweep<float><<<g, bl, 0, stream()>>>(A->coef, A->pitch, d.nx, d.ny, d.nz, A->dinv32, b->data, x->data, t,
                                                      nullptr);
      else
        k_jacobi_sweep<double><<<g, bl, 0, stream()>>>(A->coef, A->pitch, d.nx, d.ny, d.nz, A->dinv, b->data, x->data,
                                                       t, nullptr);
      FPC_LAUNCH_CHECK();
      rc = vec_copy(x->data, t, n);
      if (rc) return rc;
    }
  }
  if (rc) return rc;
  FPC_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int fpc_StructJacobiDestroy(fpc_StructSolver s) {
  if (!s) return 0;
  if (s->kind == SK_JACOBI) delete (JacobiImpl *)s->impl;
  delete s;
  return 0;
}

}  // extern "C"

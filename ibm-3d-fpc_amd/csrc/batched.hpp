// batched.hpp -- internal interface of the batched Struct solves (libfpc).
//
// A batch group holds many GMRES solvers (each with its own matrix, vectors,
// workspace and device state) and advances all of them with shared launches:
// one launch per kernel of an Arnoldi step or of a V-cycle level covers every
// system of the group.  The host runs each system's restart and convergence
// logic in lockstep with the others, so every system computes exactly what its
// own fpc_StructGMRESSolve computes.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "fpc_internal.hpp"
#include "krylov_device.hpp"

namespace fpc {

// One level of a PFMG-style hierarchy as the batched V-cycle sees it.
struct MgLevelView {
  int nx = 0, ny = 0, nz = 0;
  long long N = 0;
  int cdir = -1;
  double w = 1.0;  // relaxation weight of the level
  // what the z-marching kernels read: the hot arrays of an extruded level (rows by zm)
  long long pitch = 0;
  ZMap zm;
  const double *coef = nullptr, *dinv = nullptr, *wm = nullptr, *wp = nullptr;
  // the full arrays (one-point kernels)
  long long fpitch = 0;
  const double *fcoef = nullptr, *fdinv = nullptr, *fwm = nullptr, *fwp = nullptr;
  double *b = nullptr, *r = nullptr, *y = nullptr;
};

// The workspace and settings of a set-up GMRES solver.
struct GmresBatchView {
  int pc = 0;  // 0 none, 1 Jacobi fused into the matvec, 3 PFMG-style V-cycle
  int flexible = 0;  // Z holds kd preconditioned vectors (x += sum y_j Z_j), else one
  double *V = nullptr, *Z = nullptr, *W = nullptr, *part = nullptr;
  kry::State *S = nullptr;
  long long N = 0, vp = 0;
  int kd = 0, nred = 1, maxit = 0;
  double tol = 0.0, atol = 0.0;
  const double *dinv = nullptr;  // fused Jacobi: 1/diag at every point, or null
  fpc_StructMatrix_s *A = nullptr;
  void *pfmg = nullptr;
  int *pred = nullptr, *its = nullptr;
  double *relres = nullptr;
  bool *x0_zero = nullptr;  // the solver's zero-initial-guess hint (gmres_next_guess_is_zero)
};

// Views of set-up solvers (krylov.hip, pfmg.hip).  Errors for settings the
// batched kernels do not cover (flexible GMRES other than with the PFMG-style
// preconditioner, callback or multi-sweep preconditioners, FP32 or red/black
// V-cycles, more than one cycle).
int gmres_batch_view(fpc_StructSolver s, GmresBatchView *v);
int pfmg_batch_view(void *pfmg, std::vector<MgLevelView> &levels, const double **inv, int *ncoarse);

class BatchGroup;

// Builds a group over solvers[q] solving A x = b with the given vectors
// (data pointers of Struct vectors); all solvers must be set up.
int batch_group_create(const std::vector<fpc_StructSolver> &solvers, const std::vector<const double *> &b,
                       const std::vector<double *> &x, BatchGroup **out);
// Solves every system of the group; per system its, relative residual and
// return code (0 or FPC_ERROR_CONV) land in its/rel/err.
int batch_group_solve(BatchGroup *g, int *its, double *rel, int *err);
void batch_group_destroy(BatchGroup *g);
// The tile configuration the group's autotuner picked: z-chunk planes and the
// smallest box on z-marching tiles.
void batch_group_tiles(const BatchGroup *g, int *kz, long long *tile_min);

// Non-finite check of many device arrays at once: flags[q] = 1 when arrays[q]
// (len[q] doubles) holds a NaN or Inf.  Enqueued only; the caller reads flags
// after synchronising the stream.
class NonfiniteCheck;
int nonfinite_create(const std::vector<const double *> &arrays, const std::vector<long long> &len,
                     NonfiniteCheck **out);
int nonfinite_enqueue(NonfiniteCheck *c);
const int *nonfinite_flags(NonfiniteCheck *c);  // host (pinned), valid after a sync
void nonfinite_destroy(NonfiniteCheck *c);

}  // namespace fpc

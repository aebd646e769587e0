/*
 * fpc_struct.h -- C ABI of the MI355X Struct-stencil stack (libfpc.so).
 *
 * Drop-in for the HYPRE Struct calls and helpers of the reference
 * (JD63021/ibm-3d-fpc).  Every function returns 0 on success or an error
 * bit set (FPC_ERROR_*), mirroring HYPRE_Int error codes; the handle helpers
 * return NULL on failure.  Library-owned work runs on one HIP stream; each
 * entry point has finished when it returns, as the reference brackets the
 * corresponding HYPRE calls with cudaDeviceSynchronize (main.cu:260-351).
 * Value buffers may be host or device memory (hipMemcpyDefault resolves it);
 * the reference passes cudaMallocManaged buffers in HYPRE device mode.
 *
 * Reference interfaces replaced (file:line in the reference):
 *   CreateGrid3D / CreateStencil3D / SetMatrix3D / CreateMatrix3D /
 *   SetVector3D / GetVector3D / CreateVector3D      hypre_struct_utils.hpp:1-26
 *   MallocManaged{C,D,I,B} / CopyVecToManaged{C,D,I,B}  hypre_struct_utils.hpp:27-43
 *   HYPRE_Initialize / SetMemoryLocation / SetExecutionPolicy   main.cu:16-18
 *   HYPRE_StructGrid*, HYPRE_StructStencil*, HYPRE_StructMatrix*,
 *   HYPRE_StructVector*                              hypre_struct_utils.hpp:3-26; main.cu:396-404
 *   HYPRE_StructGMRES*                               main.cu:174-221, 284-297, 327-331, 391-394
 *   HYPRE_StructJacobi*                              main.cu:179-182, 190-193, 201-204
 *   HYPRE_StructPFMG*                                main.cu:213-220, 394
 *   HYPRE_ClearError / error codes                   common.hpp:27-43
 *   HYPRE_Finalize                                   main.cu:421
 */
#ifndef FPC_STRUCT_H
#define FPC_STRUCT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error bits (same values as HYPRE_ERROR_*). */
#define FPC_ERROR_GENERIC 1
#define FPC_ERROR_MEMORY 2
#define FPC_ERROR_ARG 4
#define FPC_ERROR_CONV 256

typedef struct fpc_StructGrid_s *fpc_StructGrid;
typedef struct fpc_StructStencil_s *fpc_StructStencil;
typedef struct fpc_StructMatrix_s *fpc_StructMatrix;
typedef struct fpc_StructVector_s *fpc_StructVector;
typedef struct fpc_StructSolver_s *fpc_StructSolver;

/* Preconditioner callback, HYPRE_PtrToStructSolverFcn (main.cu:182, 220). */
typedef int (*fpc_PtrToStructSolverFcn)(fpc_StructSolver, fpc_StructMatrix, fpc_StructVector, fpc_StructVector);

/* ---- lifecycle (HYPRE_Initialize / HYPRE_Finalize, main.cu:16, 421) ---- */
int fpc_Initialize(void);
int fpc_Finalize(void);
/* Memory location / execution policy are device-only here: accepted, no-ops (main.cu:17-18). */
int fpc_SetMemoryLocationDevice(void);
int fpc_SetExecutionPolicyDevice(void);
/* Current error bits; fpc_ClearError(bits) clears them (HYPRE_ClearError, common.hpp:39). */
int fpc_GetError(void);
int fpc_ClearError(int bits);
/* Last error text (static storage). */
const char *fpc_GetErrorMessage(void);
/* The HIP stream the library enqueues on (a hipStream_t); NULL = library default. */
int fpc_SetStream(void *hip_stream);
void *fpc_GetStream(void);

/* ---- grid (HYPRE_StructGrid*, hypre_struct_utils.hpp:1-6) ---- */
int fpc_StructGridCreate(int ndim, fpc_StructGrid *grid);
int fpc_StructGridSetExtents(fpc_StructGrid grid, const int *ilower, const int *iupper);
int fpc_StructGridAssemble(fpc_StructGrid grid);
int fpc_StructGridDestroy(fpc_StructGrid grid);

/* ---- stencil (HYPRE_StructStencil*, hypre_struct_utils.hpp:7-11) ----
 * Any ordering of the seven offsets {0, +-e_x, +-e_y, +-e_z} is accepted. */
int fpc_StructStencilCreate(int ndim, int size, fpc_StructStencil *stencil);
int fpc_StructStencilSetElement(fpc_StructStencil stencil, int entry, const int *offset);
int fpc_StructStencilDestroy(fpc_StructStencil stencil);

/* ---- matrix (HYPRE_StructMatrix*, hypre_struct_utils.hpp:12-20) ----
 * SetBoxValues copies: values are nentries per point, entry fastest, then x, y, z
 * (the AoS layout of solver_helpers.hpp:65).  Stored on the device as one
 * x-contiguous plane per stencil entry. */
int fpc_StructMatrixCreate(fpc_StructGrid grid, fpc_StructStencil stencil, fpc_StructMatrix *A);
int fpc_StructMatrixInitialize(fpc_StructMatrix A);
int fpc_StructMatrixSetBoxValues(fpc_StructMatrix A, const int *ilower, const int *iupper, int nentries,
                                 const int *entries, const double *values);
int fpc_StructMatrixGetBoxValues(fpc_StructMatrix A, const int *ilower, const int *iupper, int nentries,
                                 const int *entries, double *values);
int fpc_StructMatrixAssemble(fpc_StructMatrix A);
/* y = alpha A x + beta y (HYPRE_StructMatrixMatvec). */
int fpc_StructMatrixMatvec(double alpha, fpc_StructMatrix A, fpc_StructVector x, double beta, fpc_StructVector y);
/* Storage the solve loops read for an assembled operator: *wall_planes = m > 0
 * when its z-planes away from the walls repeat and they read a copy of m planes
 * per wall plus one interior plane, *flagged = points whose row differs from
 * that plane and is read in full; *wall_planes = 0: the full planes.
 * No reference counterpart (storage inspection). */
int fpc_StructMatrixGetExtruded(fpc_StructMatrix A, int *wall_planes, long long *flagged);
int fpc_StructMatrixDestroy(fpc_StructMatrix A);

/* ---- vector (HYPRE_StructVector*, hypre_struct_utils.hpp:21-26) ---- */
int fpc_StructVectorCreate(fpc_StructGrid grid, fpc_StructVector *x);
int fpc_StructVectorInitialize(fpc_StructVector x);
int fpc_StructVectorSetBoxValues(fpc_StructVector x, const int *ilower, const int *iupper, const double *values);
int fpc_StructVectorGetBoxValues(fpc_StructVector x, const int *ilower, const int *iupper, double *values);
int fpc_StructVectorSetConstantValues(fpc_StructVector x, double value);
int fpc_StructVectorAssemble(fpc_StructVector x);
/* Device pointer of the vector's storage (box order, x fastest); lets a caller's
 * kernels fill solver vectors in place instead of SetBoxValues copies. */
int fpc_StructVectorGetDataPointer(fpc_StructVector x, double **device_ptr);
int fpc_StructVectorDestroy(fpc_StructVector x);

/* ---- GMRES (HYPRE_StructGMRES*, main.cu:174-221) ----
 * Right-preconditioned restarted GMRES(kdim) with HYPRE's control flow:
 * stop when ||r|| <= max(atol, tol*||b||); b = 0 gives x = 0 and 0 iterations;
 * reaching maxit returns FPC_ERROR_CONV (the caller clears it, common.hpp:39). */
int fpc_StructGMRESCreate(fpc_StructSolver *solver);
int fpc_StructGMRESSetTol(fpc_StructSolver solver, double tol);
int fpc_StructGMRESSetAbsoluteTol(fpc_StructSolver solver, double atol);
int fpc_StructGMRESSetMaxIter(fpc_StructSolver solver, int maxit);
int fpc_StructGMRESSetKDim(fpc_StructSolver solver, int kdim);
int fpc_StructGMRESSetPrintLevel(fpc_StructSolver solver, int level);
int fpc_StructGMRESSetLogging(fpc_StructSolver solver, int logging);
int fpc_StructGMRESSetPrecond(fpc_StructSolver solver, fpc_PtrToStructSolverFcn precond,
                              fpc_PtrToStructSolverFcn precond_setup, fpc_StructSolver precond_solver);
int fpc_StructGMRESSetup(fpc_StructSolver solver, fpc_StructMatrix A, fpc_StructVector b, fpc_StructVector x);
int fpc_StructGMRESSolve(fpc_StructSolver solver, fpc_StructMatrix A, fpc_StructVector b, fpc_StructVector x);
int fpc_StructGMRESGetNumIterations(fpc_StructSolver solver, int *num_iterations);
int fpc_StructGMRESGetFinalRelativeResidualNorm(fpc_StructSolver solver, double *norm);
int fpc_StructGMRESDestroy(fpc_StructSolver solver);
/* Flexible GMRES: same calls; the preconditioner may change between iterations
 * (used with the FP32 PFMG).  No reference counterpart (HYPRE has FlexGMRES for
 * ParCSR only). */
int fpc_StructFlexGMRESCreate(fpc_StructSolver *solver);

/* ---- Jacobi (HYPRE_StructJacobi*, main.cu:179-182) ----
 * One sweep from a zero guess is x = b / diag(A). */
int fpc_StructJacobiCreate(fpc_StructSolver *solver);
int fpc_StructJacobiSetMaxIter(fpc_StructSolver solver, int maxit);
int fpc_StructJacobiSetTol(fpc_StructSolver solver, double tol);
int fpc_StructJacobiSetZeroGuess(fpc_StructSolver solver);
/* No HYPRE counterpart: bits = 32 keeps the preconditioner's 1/diag as stored
 * FP32 values (the mixed-precision momentum solves); the products stay FP64,
 * so the preconditioner is a fixed diagonal and plain GMRES stays exact. */
int fpc_StructJacobiSetPrecision(fpc_StructSolver solver, int bits);
int fpc_StructJacobiSetNonZeroGuess(fpc_StructSolver solver);
int fpc_StructJacobiSetup(fpc_StructSolver solver, fpc_StructMatrix A, fpc_StructVector b, fpc_StructVector x);
int fpc_StructJacobiSolve(fpc_StructSolver solver, fpc_StructMatrix A, fpc_StructVector b, fpc_StructVector x);
int fpc_StructJacobiDestroy(fpc_StructSolver solver);

/* ---- PFMG-style semicoarsening multigrid (HYPRE_StructPFMG*, main.cu:213-220) ----
 * RelaxType: 0 Jacobi, 1 weighted Jacobi (default), 2 red/black GS symmetric,
 * 3 red/black GS forward.  RAPType 0/1 both build the 7-point lumped-Galerkin
 * coarse operators.  MaxIter V-cycles per Solve (1 as a preconditioner). */
int fpc_StructPFMGCreate(fpc_StructSolver *solver);
int fpc_StructPFMGSetMaxIter(fpc_StructSolver solver, int maxit);
int fpc_StructPFMGSetTol(fpc_StructSolver solver, double tol);
int fpc_StructPFMGSetRelaxType(fpc_StructSolver solver, int relax_type);
int fpc_StructPFMGSetRAPType(fpc_StructSolver solver, int rap_type);
int fpc_StructPFMGSetNumPreRelax(fpc_StructSolver solver, int n);
int fpc_StructPFMGSetNumPostRelax(fpc_StructSolver solver, int n);
int fpc_StructPFMGSetMaxLevels(fpc_StructSolver solver, int n);
int fpc_StructPFMGSetPrintLevel(fpc_StructSolver solver, int level);
int fpc_StructPFMGSetZeroGuess(fpc_StructSolver solver);
int fpc_StructPFMGSetNonZeroGuess(fpc_StructSolver solver);
/* 64 (default) or 32: precision of the stored coarse/fine coefficients and of the
 * smoothing, restriction and prolongation.  The coarsest solve stays FP64.
 * No reference counterpart. */
int fpc_StructPFMGSetPrecision(fpc_StructSolver solver, int bits);
/* Coarsening stops at the first level with at most n points (1..4096), solved
 * by a dense inverse; 0 (default) = min(1152, max(64, N0/2048)) for a finest
 * level of N0 points.  No reference counterpart (HYPRE's PFMG has no such
 * setting); the adjoint's transposed pressure solves set 64. */
int fpc_StructPFMGSetCoarseSize(fpc_StructSolver solver, int n);
int fpc_StructPFMGSetup(fpc_StructSolver solver, fpc_StructMatrix A, fpc_StructVector b, fpc_StructVector x);
int fpc_StructPFMGSolve(fpc_StructSolver solver, fpc_StructMatrix A, fpc_StructVector b, fpc_StructVector x);
int fpc_StructPFMGGetNumIterations(fpc_StructSolver solver, int *num_iterations);
int fpc_StructPFMGGetNumLevels(fpc_StructSolver solver, int *num_levels);
/* Level l: box dims (3 ints), coarsening direction (-1 on the coarsest), relax weight. */
int fpc_StructPFMGGetLevelInfo(fpc_StructSolver solver, int level, int *dims, int *cdir, double *weight);
/* Level l's 7-point operator as stored on the device, AoS (7 per point, entry order
 * C,-x,+x,-y,+y,-z,+z), FP32 levels widened; values holds 7 * points of the level. */
int fpc_StructPFMGGetLevelMatrix(fpc_StructSolver solver, int level, double *values);
/* The same for level l's fused smoothing kernels (coefficients, 1/diag and
 * interpolation weights).  No reference counterpart. */
int fpc_StructPFMGGetLevelExtruded(fpc_StructSolver solver, int level, int *wall_planes, long long *flagged);
int fpc_StructPFMGDestroy(fpc_StructSolver solver);

/* ---- helpers with the reference signatures (hypre_struct_utils.hpp:1-26) ---- */
fpc_StructGrid fpc_CreateGrid3D(int nx, int ny, int nz);
fpc_StructStencil fpc_CreateStencil3D(void);
int fpc_SetMatrix3D(fpc_StructMatrix A, int nx, int ny, int nz, const double *vals);
fpc_StructMatrix fpc_CreateMatrix3D(fpc_StructGrid grid, fpc_StructStencil st, int nx, int ny, int nz,
                                    const double *vals);
int fpc_SetVector3D(fpc_StructVector x, int nx, int ny, int nz, const double *vals);
int fpc_GetVector3D(fpc_StructVector x, int nx, int ny, int nz, double *vals);
fpc_StructVector fpc_CreateVector3D(fpc_StructGrid grid);

/* ---- memory helpers (hypre_struct_utils.hpp:27-43) ----
 * Plain device memory (hipMalloc) plus an explicit host-to-device copy: managed
 * memory is not used on MI355X.  Free with fpc_Free. */
int fpc_MallocManagedC(double **p, size_t n);
int fpc_MallocManagedD(double **p, size_t n);
int fpc_MallocManagedI(int **p, size_t n);
int fpc_MallocManagedB(unsigned char **p, size_t n);
int fpc_CopyVecToManagedC(double **p, const double *src, size_t n);
int fpc_CopyVecToManagedD(double **p, const double *src, size_t n);
int fpc_CopyVecToManagedI(int **p, const int *src, size_t n);
int fpc_CopyVecToManagedB(unsigned char **p, const unsigned char *src, size_t n);
int fpc_Free(void *p);
/* Device-to-host / host-to-device copies for callers without a HIP runtime of their own. */
int fpc_MemcpyToHost(void *dst, const void *src, size_t bytes);
int fpc_MemcpyToDevice(void *dst, const void *src, size_t bytes);
int fpc_DeviceSynchronize(void);

/* ---- launch tuning (no reference counterpart) ----
 * "tile_min_points": boxes with at least this many points (and nx >= 32,
 * ny >= 8) use the z-marching stencil kernels (default 524288, or the
 * FPC_TILE_MIN_POINTS environment variable); "stencil_kz": z-planes per
 * workgroup of those kernels, 0 = automatic, at most 62 used (FPC_STENCIL_KZ); "z_extrude":
 * 1 (default) = operators and MG levels whose z-planes 1..nz-2 are equal keep a
 * three-plane copy that the solve loops read, 0 = off (FPC_Z_EXTRUDE); applies
 * to operators assembled and solvers set up afterwards, with identical results;
 * "mg_wide_finest": 1 (default) = an FP32 V-cycle's finest-level kernels read its
 * FP64 input and write its FP64 output directly, 0 = through FP32 copies made by
 * conversion kernels (FPC_MG_WIDE_FINEST), with identical results; "p_flexible":
 * 1 (default) = the time step's pressure GMRES keeps the preconditioned vector of
 * every Arnoldi step and forms x += sum y_j z_j, 0 = HYPRE's right-preconditioned
 * update x += M^-1 (sum y_j v_j) with one more V-cycle (FPC_P_FLEXIBLE); the same
 * iterates, the final update differing in rounding; applies to simulations created
 * afterwards; "mg_graph": 1 = V-cycles replayed as captured HIP graphs (default 0,
 * FPC_MG_GRAPH), identical results; "mg_block": workgroup size of the one-point
 * multigrid kernels, 64 (default), 128 or 256 (FPC_MG_BLOCK), identical results. */
int fpc_SetTuning(const char *key, long long value);

/* ---- kernel timing (no reference counterpart) ----
 * HIP events around the launches of the named kernel classes, on the library
 * stream: spmv_jacobi, spmv, multidot, orth_update, mg_down, mg_up, mg_restrict,
 * mg_relax, mg_residual, explicit, divergence, projection ("*" = all, NULL or ""
 * = off).  Report lines: "class launches total_ms algorithmic_bytes\n"; returns
 * the length needed including the terminating zero. */
int fpc_ProfileEnable(const char *names);
/* Time only every n-th launch of the enabled classes (n >= 1; 1 = every launch, the
 * default): the event cost then falls on 1/n of the launches. */
int fpc_ProfileSample(int every);
/* Drops the records and frees their HIP events (after the stream is synchronised):
 * recorded events left alive slow the runtime's later launches and synchronisations. */
int fpc_ProfileReset(void);
int fpc_ProfileReport(char *buf, int buflen);

#ifdef __cplusplus
}
#endif
#endif /* FPC_STRUCT_H */

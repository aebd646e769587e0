/*
 * fpc_hypre_compat.h -- the HYPRE Struct names used by the reference, mapped
 * onto the libfpc C ABI (include/fpc_struct.h).  A caller written against
 * HYPRE Struct (the reference's main.cu:16-221, 284-404 and
 * hypre_struct_utils.hpp:1-43) includes this header instead of HYPRE.h,
 * HYPRE_struct_ls.h and HYPRE_krylov.h and links -lfpc.  MPI communicators
 * are accepted and ignored (one GPU, one rank).  No new symbols: macros and
 * typedefs only.
 */
#ifndef FPC_HYPRE_COMPAT_H
#define FPC_HYPRE_COMPAT_H

#include "fpc_struct.h"

typedef int HYPRE_Int;
typedef double HYPRE_Real;
typedef double HYPRE_Complex;
typedef fpc_StructGrid HYPRE_StructGrid;
typedef fpc_StructStencil HYPRE_StructStencil;
typedef fpc_StructMatrix HYPRE_StructMatrix;
typedef fpc_StructVector HYPRE_StructVector;
typedef fpc_StructSolver HYPRE_StructSolver;
typedef fpc_PtrToStructSolverFcn HYPRE_PtrToStructSolverFcn;

#define HYPRE_ERROR_GENERIC FPC_ERROR_GENERIC
#define HYPRE_ERROR_MEMORY FPC_ERROR_MEMORY
#define HYPRE_ERROR_ARG FPC_ERROR_ARG
#define HYPRE_ERROR_CONV FPC_ERROR_CONV
#define HYPRE_MEMORY_DEVICE 1
#define HYPRE_EXEC_DEVICE 1

#define HYPRE_Initialize() fpc_Initialize()
#define HYPRE_Finalize() fpc_Finalize()
#define HYPRE_SetMemoryLocation(loc) fpc_SetMemoryLocationDevice()
#define HYPRE_SetExecutionPolicy(pol) fpc_SetExecutionPolicyDevice()
#define HYPRE_ClearError(code) fpc_ClearError(code)
#define HYPRE_GetError() fpc_GetError()

#define HYPRE_StructGridCreate(comm, ndim, grid) fpc_StructGridCreate((ndim), (grid))
#define HYPRE_StructGridSetExtents(grid, il, iu) fpc_StructGridSetExtents((grid), (il), (iu))
#define HYPRE_StructGridAssemble(grid) fpc_StructGridAssemble(grid)
#define HYPRE_StructGridDestroy(grid) fpc_StructGridDestroy(grid)

#define HYPRE_StructStencilCreate(ndim, size, st) fpc_StructStencilCreate((ndim), (size), (st))
#define HYPRE_StructStencilSetElement(st, e, off) fpc_StructStencilSetElement((st), (e), (off))
#define HYPRE_StructStencilDestroy(st) fpc_StructStencilDestroy(st)

#define HYPRE_StructMatrixCreate(comm, grid, st, A) fpc_StructMatrixCreate((grid), (st), (A))
#define HYPRE_StructMatrixInitialize(A) fpc_StructMatrixInitialize(A)
#define HYPRE_StructMatrixSetBoxValues(A, il, iu, ne, en, v) fpc_StructMatrixSetBoxValues((A), (il), (iu), (ne), (en), (v))
#define HYPRE_StructMatrixGetBoxValues(A, il, iu, ne, en, v) fpc_StructMatrixGetBoxValues((A), (il), (iu), (ne), (en), (v))
#define HYPRE_StructMatrixAssemble(A) fpc_StructMatrixAssemble(A)
#define HYPRE_StructMatrixMatvec(alpha, A, x, beta, y) fpc_StructMatrixMatvec((alpha), (A), (x), (beta), (y))
#define HYPRE_StructMatrixDestroy(A) fpc_StructMatrixDestroy(A)

#define HYPRE_StructVectorCreate(comm, grid, x) fpc_StructVectorCreate((grid), (x))
#define HYPRE_StructVectorInitialize(x) fpc_StructVectorInitialize(x)
#define HYPRE_StructVectorSetBoxValues(x, il, iu, v) fpc_StructVectorSetBoxValues((x), (il), (iu), (v))
#define HYPRE_StructVectorGetBoxValues(x, il, iu, v) fpc_StructVectorGetBoxValues((x), (il), (iu), (v))
#define HYPRE_StructVectorSetConstantValues(x, c) fpc_StructVectorSetConstantValues((x), (c))
#define HYPRE_StructVectorAssemble(x) fpc_StructVectorAssemble(x)
#define HYPRE_StructVectorDestroy(x) fpc_StructVectorDestroy(x)

#define HYPRE_StructGMRESCreate(comm, s) fpc_StructGMRESCreate(s)
#define HYPRE_StructGMRESSetTol(s, t) fpc_StructGMRESSetTol((s), (t))
#define HYPRE_StructGMRESSetAbsoluteTol(s, t) fpc_StructGMRESSetAbsoluteTol((s), (t))
#define HYPRE_StructGMRESSetMaxIter(s, n) fpc_StructGMRESSetMaxIter((s), (n))
#define HYPRE_StructGMRESSetKDim(s, n) fpc_StructGMRESSetKDim((s), (n))
#define HYPRE_StructGMRESSetPrintLevel(s, n) fpc_StructGMRESSetPrintLevel((s), (n))
#define HYPRE_StructGMRESSetLogging(s, n) fpc_StructGMRESSetLogging((s), (n))
#define HYPRE_StructGMRESSetPrecond(s, f, fs, pc) fpc_StructGMRESSetPrecond((s), (f), (fs), (pc))
#define HYPRE_StructGMRESSetup(s, A, b, x) fpc_StructGMRESSetup((s), (A), (b), (x))
#define HYPRE_StructGMRESSolve(s, A, b, x) fpc_StructGMRESSolve((s), (A), (b), (x))
#define HYPRE_StructGMRESGetNumIterations(s, n) fpc_StructGMRESGetNumIterations((s), (n))
#define HYPRE_StructGMRESGetFinalRelativeResidualNorm(s, r) fpc_StructGMRESGetFinalRelativeResidualNorm((s), (r))
#define HYPRE_StructGMRESDestroy(s) fpc_StructGMRESDestroy(s)

#define HYPRE_StructJacobiCreate(comm, s) fpc_StructJacobiCreate(s)
#define HYPRE_StructJacobiSetMaxIter(s, n) fpc_StructJacobiSetMaxIter((s), (n))
#define HYPRE_StructJacobiSetTol(s, t) fpc_StructJacobiSetTol((s), (t))
#define HYPRE_StructJacobiSetZeroGuess(s) fpc_StructJacobiSetZeroGuess(s)
#define HYPRE_StructJacobiSetNonZeroGuess(s) fpc_StructJacobiSetNonZeroGuess(s)
#define HYPRE_StructJacobiSetup fpc_StructJacobiSetup
#define HYPRE_StructJacobiSolve fpc_StructJacobiSolve
#define HYPRE_StructJacobiDestroy(s) fpc_StructJacobiDestroy(s)

#define HYPRE_StructPFMGCreate(comm, s) fpc_StructPFMGCreate(s)
#define HYPRE_StructPFMGSetMaxIter(s, n) fpc_StructPFMGSetMaxIter((s), (n))
#define HYPRE_StructPFMGSetTol(s, t) fpc_StructPFMGSetTol((s), (t))
#define HYPRE_StructPFMGSetRelaxType(s, t) fpc_StructPFMGSetRelaxType((s), (t))
#define HYPRE_StructPFMGSetRAPType(s, t) fpc_StructPFMGSetRAPType((s), (t))
#define HYPRE_StructPFMGSetNumPreRelax(s, n) fpc_StructPFMGSetNumPreRelax((s), (n))
#define HYPRE_StructPFMGSetNumPostRelax(s, n) fpc_StructPFMGSetNumPostRelax((s), (n))
#define HYPRE_StructPFMGSetMaxLevels(s, n) fpc_StructPFMGSetMaxLevels((s), (n))
#define HYPRE_StructPFMGSetPrintLevel(s, n) fpc_StructPFMGSetPrintLevel((s), (n))
#define HYPRE_StructPFMGSetZeroGuess(s) fpc_StructPFMGSetZeroGuess(s)
#define HYPRE_StructPFMGSetNonZeroGuess(s) fpc_StructPFMGSetNonZeroGuess(s)
#define HYPRE_StructPFMGSetup fpc_StructPFMGSetup
#define HYPRE_StructPFMGSolve fpc_StructPFMGSolve
#define HYPRE_StructPFMGGetNumIterations(s, n) fpc_StructPFMGGetNumIterations((s), (n))
#define HYPRE_StructPFMGDestroy(s) fpc_StructPFMGDestroy(s)

/* hypre_struct_utils.hpp helper names */
#define CreateGrid3D fpc_CreateGrid3D
#define CreateStencil3D fpc_CreateStencil3D
#define SetMatrix3D fpc_SetMatrix3D
#define CreateMatrix3D fpc_CreateMatrix3D
#define SetVector3D fpc_SetVector3D
#define GetVector3D fpc_GetVector3D
#define CreateVector3D fpc_CreateVector3D

#ifdef __cplusplus
#include <vector>
/* hypre_struct_utils.hpp:27-43 take std::vector sources; these inline forms keep
 * those call sites unchanged.  The buffers are device memory (hipMalloc): host
 * code reads them back with fpc_MemcpyToHost. */
static inline void MallocManagedC(double **p, size_t n) { (void)fpc_MallocManagedC(p, n); }
static inline void MallocManagedD(double **p, size_t n) { (void)fpc_MallocManagedD(p, n); }
static inline void MallocManagedI(int **p, size_t n) { (void)fpc_MallocManagedI(p, n); }
static inline void MallocManagedB(unsigned char **p, size_t n) { (void)fpc_MallocManagedB(p, n); }
static inline void CopyVecToManagedC(double **p, const std::vector<double> &src) {
  (void)fpc_CopyVecToManagedC(p, src.data(), src.size());
}
static inline void CopyVecToManagedD(double **p, const std::vector<double> &src) {
  (void)fpc_CopyVecToManagedD(p, src.data(), src.size());
}
static inline void CopyVecToManagedI(int **p, const std::vector<int> &src) {
  (void)fpc_CopyVecToManagedI(p, src.data(), src.size());
}
static inline void CopyVecToManagedB(unsigned char **p, const std::vector<char> &src) {
  (void)fpc_CopyVecToManagedB(p, (const unsigned char *)src.data(), src.size());
}
#endif

#endif /* FPC_HYPRE_COMPAT_H */

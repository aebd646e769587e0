// fpc_internal.hpp -- internal definitions shared by the libfpc HIP sources.
//
// Layout in HBM: a Struct box of n = (nx, ny, nz) points is stored x fastest,
// then y, then z (common.hpp:230-238 of the reference).  A 7-point matrix is
// stored as seven x-contiguous planes ("SoA"), canonical slot order
// C, -x, +x, -y, +y, -z, +z (the reference stencil order, hypre_struct_utils.hpp:9),
// each plane padded to a 256-byte multiple so every plane starts aligned.
//
// An operator extruded along z (a full-height obstacle between z walls: the
// planes away from the walls repeat) also keeps a "hot" copy for the stencil
// kernels of the solve loops: m planes at each wall, one interior plane, then
// all rows again.  Points whose row differs from the interior plane (rounding
// in the assembly, a boundary layer deeper than m) are flagged and read their
// own row from the full part, so results are those of the full planes bit for
// bit while the coefficient traffic is about 2m+1 planes instead of nz.  Every
// other reader keeps the full planes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "fpc_struct.h"

namespace fpc {

void set_error(int bits, const char *fmt, ...);
int error_bits();
hipStream_t stream();
// Whether stream() has a stream without creating one (a handle's scoped stream, a stream
// adopted by fpc_SetStream, or the library stream made by fpc_Initialize).
bool stream_ready();
void ensure_init();
// Frees the calling thread's scratch buffers (worker threads, before they exit).
void release_thread_state();

// Routes this thread's library calls to stream s for the scope's lifetime
// (each simulation handle runs on a stream of its own).
class StreamScope {
 public:
  explicit StreamScope(hipStream_t s);
  ~StreamScope();
  StreamScope(const StreamScope &) = delete;
  StreamScope &operator=(const StreamScope &) = delete;

 private:
  hipStream_t prev_;
};

#define FPC_HIP(call)                                                                                  \
  do {                                                                                                 \
    hipError_t e_ = (call);                                                                            \
    if (e_ != hipSuccess) {                                                                            \
      ::fpc::set_error(FPC_ERROR_GENERIC, "%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
      return FPC_ERROR_GENERIC;                                                                        \
    }                                                                                                  \
  } while (0)

#define FPC_CHECK_ARG(cond, msg)                                          \
  do {                                                                    \
    if (!(cond)) {                                                        \
      ::fpc::set_error(FPC_ERROR_ARG, "%s: %s", __func__, msg);          \
      return FPC_ERROR_ARG;                                               \
    }                                                                     \
  } while (0)

// Kernel timing with HIP events on the library stream (fpc_ProfileEnable); for
// one calling thread at a time.
bool prof_enabled();
bool prof_on(const char *cls);
void prof_begin(const char *cls);
void prof_end(const char *cls, double bytes);
// Launches enqueued ahead of a device-side stop flag may return at once; they are
// tagged with the Arnoldi step they belong to and its solver (owner), and dropped once
// the host knows how many steps ran (that owner's records from index `from` with
// tag >= ran; solves advanced in lockstep interleave their records).
size_t prof_mark();
void prof_tag(int tag, const void *owner);
void prof_drop_skipped(size_t from, int ran, const void *owner);
#define FPC_PROF_BEGIN(cls)                    \
  const bool prof_on_ = ::fpc::prof_on(cls);   \
  if (prof_on_) ::fpc::prof_begin(cls)
#define FPC_PROF_END(cls, bytes) \
  if (prof_on_) ::fpc::prof_end(cls, (double)(bytes))

// Launch check: reports a failed launch as an error return.
#define FPC_LAUNCH_CHECK() FPC_HIP(hipGetLastError())

constexpr int kSlots = 7;
constexpr int kMaxKDim = 128;       // largest GMRES restart length supported
constexpr int kRedBlocks = 1024;    // most blocks of the deterministic reductions

// Blocks of a deterministic reduction over n points: one per 2048 points, at
// most kRedBlocks.  The partition depends on n only, so every reduction of a
// vector (alone or in a batch) sums the same partials in the same order.
inline int red_blocks(long long n) {
  long long b = (n + 2047) / 2048;
  return (int)(b < 1 ? 1 : (b > kRedBlocks ? kRedBlocks : b));
}

inline long long pad256(long long n, int elsize) {
  long long per = 256 / elsize;
  return (n + per - 1) / per * per;
}

// Box geometry for launches: 2-D blocks over (x, y), grid z = box z.
struct Dims {
  int nx = 0, ny = 0, nz = 0;
  long long N = 0;
};

// threads: workgroup size (a power of two, 64..256).  Smaller workgroups spread a
// small box over more CUs.
inline void launch_shape(const Dims &d, dim3 &grid, dim3 &block, int threads = 256) {
  int tx = 64;
  while (tx > 1 && tx / 2 >= d.nx) tx /= 2;
  int ty = threads / tx;
  if (ty < 1) ty = 1;
  while (ty > 1 && ty / 2 >= d.ny) ty /= 2;
  block = dim3(tx, ty, 1);
  grid = dim3((d.nx + tx - 1) / tx, (d.ny + ty - 1) / ty, d.nz);
}

// v unchanged, but opaque to the optimiser: an index chosen by a select stays one
// address (x[in ? t + o : t] is not split into a branch around a second load).
template <typename T>
__device__ inline T opaque(T v) {
  asm("" : "+v"(v));
  return v;
}

// No instruction: v is computed before what follows, so loads it depends on are
// not sunk below a later early return (a status test after the loads).
template <typename T>
__device__ inline void pin(T v) {
  asm volatile("" ::"v"(v));
}

// Each value computed here (one wait for all the loads they come from): code after
// this uses them under its own conditions without the loads being sunk into those.
template <typename... T>
__device__ inline void pin_all(const T &...v) {
  (pin(v), ...);
}

// (i, j, k) += (si, sj, sk) with carries; si < nx and sj < ny, so one carry per direction at most.
__device__ inline void ijk_advance(int &i, int &j, int &k, int si, int sj, int sk, int nx, int ny) {
  i += si;
  j += sj;
  k += sk;
  if (i >= nx) { i -= nx; ++j; }
  if (j >= ny) { j -= ny; ++k; }
}

// A uniform value (held in scalar registers) made opaque in the same way as opaque().
template <typename T>
__device__ inline T opaque_s(T v) {
  asm("" : "+s"(v));
  return v;
}

// The flag word read by points of a map without flagged points.
static __device__ const unsigned long long kNoFlags[1] = {0ull};

// Row of the stored planes that holds point t (on z-plane k).  pl == 0: full
// storage, row t.  Otherwise the hot copy: a flagged point reads row xoff + t;
// any other reads compact plane kc -- k for the m bottom planes, m for the
// interior, m+1 .. 2m for the m top planes -- at row t + (kc - k) * pl.  The
// flags are bits per column: bit (k & 63) of xb[(k >> 6) * pl + (t - k pl)].
struct ZMap {
  long long pl = 0;    // nx * ny of the box, or 0
  int nz = 0, m = 0;   // box depth, wall planes kept per side
  long long xoff = 0;  // first row of the full part
  const unsigned long long *xb = nullptr;  // flag words (null: no flagged point)
  __host__ __device__ inline long long hot(long long t, int k) const {
    const int kc = (k < m) ? k : ((k >= nz - m) ? k - nz + 2 * m + 1 : m);
    return t + (long long)(kc - k) * pl;
  }
  __host__ __device__ inline long long operator()(long long t, int k) const {
    if (!pl) return t;
    if (xb && ((xb[(long long)(k >> 6) * pl + (t - (long long)k * pl)] >> (k & 63)) & 1ull)) return xoff + t;
    return hot(t, k);
  }
  // The row when the caller knows the point's flag f.
  __host__ __device__ inline long long row(long long t, int k, int f) const {
    if (!pl) return t;
    return f ? xoff + t : hot(t, k);
  }
  // Flag word w of column c (j nx + i); 0 without flags.
  __device__ inline unsigned long long word(long long c, int w) const {
    return (pl && xb) ? xb[(long long)w * pl + c] : 0ull;
  }
  // The rows of n points at once, in two steps so that other loads can go between
  // them: flags() issues every flag load (box-wide tests once), rows() uses them.
  // No branch: without flags every lane reads one zero word, so the loads' wait can
  // move down to rows() behind whatever the caller issues in between.
  template <int n>
  __device__ inline void flags(const long long *t, const int *k, unsigned char *f) const {
    const bool on = pl && xb;
    typedef const __attribute__((address_space(1))) unsigned long long *gptr;  // a global load, not flat
    const gptr p = (gptr)opaque_s(on ? xb : kNoFlags);                          // not folded to a zero
    for (int q = 0; q < n; ++q)
      f[q] = (unsigned char)((p[on ? (long long)(k[q] >> 6) * pl + (t[q] - (long long)k[q] * pl) : 0] >> (k[q] & 63)) & 1ull);
  }
  template <int n>
  __device__ inline void rows(const long long *t, const int *k, const unsigned char *f, long long *u) const {
    if (!pl) {
      for (int q = 0; q < n; ++q) u[q] = t[q];
      return;
    }
    for (int q = 0; q < n; ++q) {
      const long long h = hot(t[q], k[q]), x = xoff + t[q];
      u[q] = f[q] ? x : h;
    }
  }
};

// The full-storage map as a type: row t, no flags, no branches (the one-point
// kernels, whose neighbour loads must issue together).
struct FullMap {
  __host__ __device__ inline long long operator()(long long t, int) const { return t; }
  template <int n>
  __device__ inline void flags(const long long *, const int *, unsigned char *) const {}
  __device__ inline long long row(long long t, int, int) const { return t; }
  template <int n>
  __device__ inline void rows(const long long *t, const int *, const unsigned char *, long long *u) const {
    for (int q = 0; q < n; ++q) u[q] = t[q];
  }
};

template <typename T>
struct Stencil {
  const T *c;        // 7 planes
  long long pitch;   // elements between planes
  int nx, ny, nz;
};

// s + a v (one fused multiply-add) when `in`, else s: the fma always runs with a zeroed
// coefficient, so a and v are used on every path and their loads cannot be sunk into a
// branch of their own (a wait per neighbour).  Bitwise s + a v, or s, for finite v
// (s = -0 becomes +0, which only changes the sign of a zero result).
__device__ inline double fma_if(bool in, double s, double a, double v) { return __builtin_fma(in ? a : 0.0, v, s); }
__device__ inline float fma_if(bool in, float s, float a, float v) { return __builtin_fmaf(in ? a : 0.0f, v, s); }

// Deterministic block reduction of one double per thread (wave64 shuffles + LDS).
template <int BS>
__device__ inline double block_sum(double v, double *lds) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < BS / 64; ++w) r += lds[w];
  }
  return r;  // valid in thread 0
}

}  // namespace fpc

// ---------------------------------------------------------------- handles
struct fpc_StructGrid_s {
  int ndim = 3;
  int il[3] = {0, 0, 0}, iu[3] = {-1, -1, -1};
  int n[3] = {0, 0, 0};
  long long N = 0;
  bool extents = false, assembled = false;
};

struct fpc_StructStencil_s {
  int ndim = 3, size = 0;
  int off[fpc::kSlots][3] = {};
  int set_mask = 0;
};

struct fpc_StructMatrix_s {
  fpc_StructGrid_s grid;
  int slot_of_entry[fpc::kSlots] = {0, 1, 2, 3, 4, 5, 6};
  int nentries = 7;
  long long pitch = 0;
  double *coef = nullptr;    // device, 7 * pitch
  double *dinv = nullptr;    // device, 1 / diagonal (Jacobi)
  float *coef32 = nullptr;   // device, FP32 copy (mixed-precision paths)
  float *dinv32 = nullptr;
  // hot copies of an extruded operator (zpitch apart, rows by zmap; zmap.pl != 0 when present)
  double *zcoef = nullptr, *zdinv = nullptr;
  float *zcoef32 = nullptr, *zdinv32 = nullptr;
  unsigned long long *zxf = nullptr;  // zmap.xb
  long long zpitch = 0, zexc = 0;  // zexc: flagged points
  fpc::ZMap zmap;
  bool initialized = false, assembled = false;
  unsigned long long version = 0;  // bumped whenever values change
};

struct fpc_StructVector_s {
  fpc_StructGrid_s grid;
  double *data = nullptr;
  bool initialized = false;
};

struct fpc_StructSolver_s {
  int kind = 0;
  void *impl = nullptr;
};

namespace fpc {

enum SolverKind { SK_NONE = 0, SK_GMRES = 1, SK_FGMRES = 2, SK_JACOBI = 3, SK_PFMG = 4 };

// Native preconditioners used inside GMRES (jacobi.hip, pfmg.hip).
struct JacobiImpl {
  int bits = 64;  // 32: 1/diag stored in FP32 (the products stay FP64)
  int maxit = 1;
  double tol = 0.0;
  bool zero_guess = false;
  fpc_StructMatrix_s *A = nullptr;
};
// x = M^-1 b with a zero initial guess (maxit sweeps); skip: device flag, work is skipped when *skip != 0.
int jacobi_apply(JacobiImpl *J, const double *b, double *x, const int *skip);
int pfmg_setup(void *pf, fpc_StructMatrix_s *A);
int pfmg_apply(void *pf, const double *b, double *x, const int *skip);

inline Dims dims_of(const fpc_StructGrid_s &g) {
  Dims d;
  d.nx = g.n[0]; d.ny = g.n[1]; d.nz = g.n[2]; d.N = g.N;
  return d;
}

// Vector kernels used by several solvers (struct_mv.hip).
int vec_copy(double *dst, const double *src, long long n);
int vec_zero(double *dst, long long n);
// x += y
int axpy_inplace(double *x, const double *y, long long n);
// r = b - A x (FP64), writes partial sums of r*r per block into part (red_blocks(N) of them)
int residual_norm_partials(const fpc_StructMatrix_s *A, const double *b, const double *x, double *r, double *part);
// sum of nparts partials into *out (device)
int reduce_partials(const double *part, int nparts, double *out);
int dot_partials(const double *a, const double *b, long long n, double *part);
// r = b with the partial sums of b*b that residual_norm_partials gives for x = 0 (bit for bit)
int copy_norm_partials(const double *b, double *r, long long n, double *part);
int matvec(double alpha, const fpc_StructMatrix_s *A, const double *x, double beta, double *y);
int ensure_dinv(fpc_StructMatrix_s *A);
int ensure_fp32(fpc_StructMatrix_s *A);

// The next solve of GMRES solver s starts from x = 0 (its caller has just written a zero
// initial guess): its initial residual is then b itself, copied instead of computed with
// a matvec (the same bits).  Cleared by that solve.
void gmres_next_guess_is_zero(fpc_StructSolver s);
void gmres_clear_guess_hint(fpc_StructSolver s);

// n GMRES solves (fpc_StructGMRESSolve each) advanced in lockstep with one host sync
// per round for all of them (krylov.hip); rc[q] is what solve q alone would return.
// Returns nonzero (and abandons the others) when a solve fails outside convergence.
int gmres_solve_many(const fpc_StructSolver *s, const fpc_StructMatrix *A, const fpc_StructVector *b,
                     const fpc_StructVector *x, int n, int *rc);

// The planes the solve loops' stencil kernels read: the extruded copy when the
// operator has one, else the full planes (ensure_dinv / ensure_fp32 first).
template <typename T>
struct Planes {
  const T *c = nullptr, *dinv = nullptr;
  long long pitch = 0;
  ZMap zm;
};
inline Planes<double> planes64(const fpc_StructMatrix_s *A) {
  Planes<double> p;
  if (A->zmap.pl) {
    p.c = A->zcoef; p.dinv = A->zdinv; p.pitch = A->zpitch; p.zm = A->zmap;
  } else {
    p.c = A->coef; p.dinv = A->dinv; p.pitch = A->pitch;
  }
  return p;
}
inline Planes<float> planes32(const fpc_StructMatrix_s *A) {
  Planes<float> p;
  if (A->zmap.pl) {
    p.c = A->zcoef32; p.dinv = A->zdinv32; p.pitch = A->zpitch; p.zm = A->zmap;
  } else {
    p.c = A->coef32; p.dinv = A->dinv32; p.pitch = A->pitch;
  }
  return p;
}
// The 1/diag the fused Jacobi matvec reads: in the storage planes64 / planes32
// use (hot), or the full plane.
template <typename DT>
inline const DT *jacobi_dinv(const fpc_StructMatrix_s *A, bool hot = true);
template <>
inline const double *jacobi_dinv<double>(const fpc_StructMatrix_s *A, bool hot) {
  return (hot && A->zmap.pl) ? A->zdinv : A->dinv;
}
template <>
inline const float *jacobi_dinv<float>(const fpc_StructMatrix_s *A, bool hot) {
  return (hot && A->zmap.pl) ? A->zdinv32 : A->dinv32;
}
// Coefficient rows a stencil pass reads: the compact planes and the flagged
// points' own rows of a hot copy, or every row.
inline double stored_rows(const ZMap &zm, long long N, long long nexc) {
  return zm.pl ? (double)(2 * zm.m + 1) * (double)zm.pl + (double)nexc : (double)N;
}
// Whether new operators and MG levels get extruded copies ("z_extrude" tuning key).
bool z_extrude_enabled();
// Device scratch owned by the library (grows on demand, freed at Finalize).
double *scratch_doubles(size_t n, int which);

}  // namespace fpc

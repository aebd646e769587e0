// struct_mv.hip -- Struct grid/stencil/matrix/vector objects, the 7-point
// matvec and the deterministic reductions (libfpc).
//
// Replaces the reference's HYPRE Struct object calls behind the helper
// signatures of hypre_struct_utils.hpp:1-43 (see include/fpc_struct.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>

#include "extrude.hpp"
#include "fpc_internal.hpp"
#include "krylov_device.hpp"
#include "stencil_tile.hpp"

namespace fpc {

namespace {
// Library-wide default stream (fpc_SetStream), created on first use.
std::mutex g_mu;
std::atomic<bool> g_init{false};
hipStream_t g_stream = nullptr;
bool g_own_stream = false;

// Per calling thread: error bits and message (HYPRE keeps one global error
// state; here each host thread has its own), the stream a handle scoped for
// its calls, and scratch buffers, so handles can be driven from several
// threads at once.
constexpr int kScratchSlots = 8;
struct ThreadState {
  int error = 0;
  char msg[512] = "";
  hipStream_t scoped = nullptr;
  double *scratch[kScratchSlots] = {};
  size_t scratch_n[kScratchSlots] = {};
};
thread_local ThreadState t_state;
}  // namespace

void set_error(int bits, const char *fmt, ...) {
  t_state.error |= bits;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(t_state.msg, sizeof(t_state.msg), fmt, ap);
  va_end(ap);
}

int error_bits() { return t_state.error; }

void ensure_init() {
  if (g_init.load(std::memory_order_acquire)) return;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_init.load(std::memory_order_relaxed)) return;
  if (!g_stream) {
    if (hipStreamCreate(&g_stream) != hipSuccess) {
      set_error(FPC_ERROR_GENERIC, "hipStreamCreate failed");
      g_stream = nullptr;
    } else {
      g_own_stream = true;
    }
  }
  g_init.store(true, std::memory_order_release);
}

hipStream_t stream() {
  if (t_state.scoped) return t_state.scoped;
  ensure_init();
  return g_stream;
}

bool stream_ready() { return t_state.scoped != nullptr || g_init.load(std::memory_order_acquire); }

StreamScope::StreamScope(hipStream_t s) : prev_(t_state.scoped) { t_state.scoped = s; }
StreamScope::~StreamScope() { t_state.scoped = prev_; }

double *scratch_doubles(size_t n, int which) {
  if (which < 0 || which >= kScratchSlots) return nullptr;
  ThreadState &ts = t_state;
  if (ts.scratch_n[which] < n) {
    if (ts.scratch[which]) {
      (void)hipStreamSynchronize(stream());  // the old buffer may still be in use
      (void)hipFree(ts.scratch[which]);
    }
    ts.scratch[which] = nullptr;
    ts.scratch_n[which] = 0;
    void *p = nullptr;
    if (hipMalloc(&p, n * sizeof(double)) != hipSuccess) {
      set_error(FPC_ERROR_MEMORY, "scratch allocation of %zu doubles failed", n);
      return nullptr;
    }
    ts.scratch[which] = (double *)p;
    ts.scratch_n[which] = n;
  }
  return ts.scratch[which];
}

void release_thread_state() {
  ThreadState &ts = t_state;
  for (int w = 0; w < kScratchSlots; ++w) {
    if (ts.scratch[w]) (void)hipFree(ts.scratch[w]);
    ts.scratch[w] = nullptr;
    ts.scratch_n[w] = 0;
  }
}

// ---------------------------------------------------------------- kernel timing
namespace {
struct ProfRec {
  std::string cls;
  double bytes;
  hipEvent_t e0, e1;
  int tag;
  const void *owner;
  bool dropped;
};
// The records are process-wide and unsynchronised: timing is for one calling
// thread at a time (fpc_ensemble_step refuses a multi-threaded pool while it is
// on).  With timing off nothing here touches shared state.
std::string g_prof_filter;
std::atomic<bool> g_prof{false};
std::vector<ProfRec> g_prof_recs;
std::vector<std::pair<hipEvent_t, hipEvent_t>> g_prof_pool;
size_t g_prof_used = 0;
hipEvent_t g_prof_open = nullptr;
int g_prof_every = 1;            // time every n-th launch of the enabled classes
unsigned long long g_prof_seen = 0;
thread_local int t_prof_tag = -1;
thread_local const void *t_prof_owner = nullptr;
}  // namespace

bool prof_enabled() { return g_prof.load(std::memory_order_relaxed); }

size_t prof_mark() { return prof_enabled() ? g_prof_recs.size() : 0; }

void prof_tag(int tag, const void *owner) {
  if (!prof_enabled()) return;
  t_prof_tag = tag;
  t_prof_owner = owner;
}

void prof_drop_skipped(size_t from, int ran, const void *owner) {
  if (!prof_enabled()) return;
  for (size_t q = from; q < g_prof_recs.size(); ++q)
    if (g_prof_recs[q].owner == owner && g_prof_recs[q].tag >= ran) g_prof_recs[q].dropped = true;
}

// Whether class cls is enabled, remembered per class name (the names are string literals,
// so their addresses identify them; cleared when the filter changes): the test runs at
// every launch of every class while timing is on.
std::vector<std::pair<const char *, bool>> g_prof_cls;
static bool prof_class_on(const char *cls) {
  if (g_prof_filter == "*") return true;
  for (const auto &e : g_prof_cls)
    if (e.first == cls) return e.second;
  std::string f = "," + g_prof_filter + ",", c = "," + std::string(cls) + ",";
  const bool on = f.find(c) != std::string::npos;
  g_prof_cls.emplace_back(cls, on);
  return on;
}

bool prof_on(const char *cls) {
  if (!prof_enabled()) return false;
  if (!prof_class_on(cls)) return false;
  if (g_prof_every <= 1) return true;
  return (g_prof_seen++ % (unsigned long long)g_prof_every) == 0;
}

void prof_begin(const char *cls) {
  (void)cls;
  if (g_prof_used == g_prof_pool.size()) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
    g_prof_pool.emplace_back(a, b);
  }
  g_prof_open = g_prof_pool[g_prof_used].first;
  (void)hipEventRecord(g_prof_open, stream());
}

void prof_end(const char *cls, double bytes) {
  if (!g_prof_open || g_prof_used >= g_prof_pool.size()) return;
  auto &pr = g_prof_pool[g_prof_used++];
  (void)hipEventRecord(pr.second, stream());
  g_prof_recs.push_back({cls, bytes, pr.first, pr.second, t_prof_tag, t_prof_owner, false});
  g_prof_open = nullptr;
}

// ---------------------------------------------------------------- kernels
namespace {

// Scatter AoS values of a sub-box into the SoA planes.
__global__ void k_set_matrix_box(double *coef, long long pitch, int gnx, int gny, int ox, int oy, int oz, int ex,
                                 int ey, int ez, int nent, const int *slots, const double *vals, int getmode,
                                 double *outvals) {
  long long npts = (long long)ex * ey * ez;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < npts; t += (long long)gridDim.x * blockDim.x) {
    int i = (int)(t % ex);
    long long q = t / ex;
    int j = (int)(q % ey);
    int k = (int)(q / ey);
    long long r = ((long long)(k + oz) * gny + (j + oy)) * gnx + (i + ox);
    for (int e = 0; e < nent; ++e) {
      int s = slots[e];
      if (getmode) outvals[t * nent + e] = coef[s * pitch + r];
      else coef[s * pitch + r] = vals[t * nent + e];
    }
  }
}

__global__ void k_vector_box(double *data, int gnx, int gny, int ox, int oy, int oz, int ex, int ey, int ez,
                             const double *in, double *out) {
  long long npts = (long long)ex * ey * ez;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < npts; t += (long long)gridDim.x * blockDim.x) {
    int i = (int)(t % ex);
    long long q = t / ex;
    int j = (int)(q % ey);
    int k = (int)(q / ey);
    long long r = ((long long)(k + oz) * gny + (j + oy)) * gnx + (i + ox);
    if (in) data[r] = in[t];
    else out[t] = data[r];
  }
}

__global__ void k_fill(double *x, long long n, double v) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) x[t] = v;
}

__global__ void k_axpy1(double *x, const double *y, long long n) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) x[t] += y[t];
}

__global__ void k_copy(double *d, const double *s, long long n) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) d[t] = s[t];
}

template <typename T>
__global__ void k_dinv(const T *coef, T *dinv, long long n) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    T d = coef[t];
    dinv[t] = (d != (T)0) ? (T)1 / d : (T)0;
  }
}

__global__ void k_to_fp32(const double *s, float *d, long long n) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) d[t] = (float)s[t];
}

// y = alpha * A x + beta * y, one thread per box point.
__global__ void k_matvec(const double *__restrict__ c, long long pitch, int nx, int ny, int nz, double alpha,
                         const double *__restrict__ x, double beta, double *__restrict__ y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  if (i >= nx || j >= ny) return;
  long long r = ((long long)k * ny + j) * nx + i;
  long long pl = (long long)nx * ny;
  double s = c[r] * x[r];
  if (i > 0) s += c[pitch + r] * x[r - 1];
  if (i < nx - 1) s += c[2 * pitch + r] * x[r + 1];
  if (j > 0) s += c[3 * pitch + r] * x[r - nx];
  if (j < ny - 1) s += c[4 * pitch + r] * x[r + nx];
  if (k > 0) s += c[5 * pitch + r] * x[r - pl];
  if (k < nz - 1) s += c[6 * pitch + r] * x[r + pl];
  double out = alpha * s;
  if (beta != 0.0) out += beta * y[r];
  y[r] = out;
}

// r = b - A x with per-block partial sums of r^2.
template <class M>
__global__ void __launch_bounds__(256) k_residual(const double *__restrict__ c, long long pitch, M zm, int nx, int ny,
                                                  int nz, const double *__restrict__ b, const double *__restrict__ x,
                                                  double *__restrict__ r, double *__restrict__ part) {
  kry::residual_body(c, pitch, zm, nx, ny, nz, b, x, r, part, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(256) k_dot(const double *__restrict__ a, const double *__restrict__ b, long long n,
                                             double *__restrict__ part) {
  kry::dot_body(a, b, n, part, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(256) k_copy_dot(const double *__restrict__ b, double *__restrict__ r, long long n,
                                                  double *__restrict__ part) {
  kry::copy_dot_body(b, r, n, part, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(1024) k_reduce(const double *__restrict__ part, int n, double *__restrict__ out) {
  kry::reduce_body(part, n, out);
}

unsigned blocks_for(long long n) {
  long long b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > 65536) b = 65536;
  return (unsigned)b;
}

}  // namespace

int vec_copy(double *dst, const double *src, long long n) {
  if (n <= 0) return 0;
  k_copy<<<blocks_for(n), 256, 0, stream()>>>(dst, src, n);
  FPC_LAUNCH_CHECK();
  return 0;
}

int axpy_inplace(double *x, const double *y, long long n) {
  if (n <= 0) return 0;
  k_axpy1<<<blocks_for(n), 256, 0, stream()>>>(x, y, n);
  FPC_LAUNCH_CHECK();
  return 0;
}

int vec_zero(double *dst, long long n) {
  if (n <= 0) return 0;
  k_fill<<<blocks_for(n), 256, 0, stream()>>>(dst, n, 0.0);
  FPC_LAUNCH_CHECK();
  return 0;
}

int matvec(double alpha, const fpc_StructMatrix_s *A, const double *x, double beta, double *y) {
  Dims d = dims_of(A->grid);
  if (d.N == 0) return 0;
  dim3 g, b;
  launch_shape(d, g, b);
  k_matvec<<<g, b, 0, stream()>>>(A->coef, A->pitch, d.nx, d.ny, d.nz, alpha, x, beta, y);
  FPC_LAUNCH_CHECK();
  return 0;
}

int residual_norm_partials(const fpc_StructMatrix_s *A, const double *b, const double *x, double *r, double *part) {
  Dims d = dims_of(A->grid);
  // the hot copy on boxes the z-marching kernels take; latency-bound small boxes the full planes
  if (A->zmap.pl && tile_worthwhile(d)) {
    const Planes<double> P = planes64(A);
    k_residual<ZMap><<<red_blocks(d.N), 256, 0, stream()>>>(P.c, P.pitch, P.zm, d.nx, d.ny, d.nz, b, x, r, part);
  } else {
    k_residual<FullMap><<<red_blocks(d.N), 256, 0, stream()>>>(A->coef, A->pitch, FullMap(), d.nx, d.ny, d.nz, b, x, r,
                                                               part);
  }
  FPC_LAUNCH_CHECK();
  return 0;
}

int copy_norm_partials(const double *b, double *r, long long n, double *part) {
  k_copy_dot<<<red_blocks(n), 256, 0, stream()>>>(b, r, n, part);
  FPC_LAUNCH_CHECK();
  return 0;
}

int dot_partials(const double *a, const double *b, long long n, double *part) {
  k_dot<<<red_blocks(n), 256, 0, stream()>>>(a, b, n, part);
  FPC_LAUNCH_CHECK();
  return 0;
}

int reduce_partials(const double *part, int nparts, double *out) {
  k_reduce<<<1, 1024, 0, stream()>>>(part, nparts, out);
  FPC_LAUNCH_CHECK();
  return 0;
}

bool z_extrude_enabled() { return tile_tuning().z_extrude != 0; }

namespace {

// FP32 hot copies of an extruded operator, once it has FP32 planes: the same
// conversions as coef32 / dinv32, over the hot layout.
int pack_extruded32(fpc_StructMatrix_s *A) {
  if (!A->zmap.pl || !A->coef32 || A->zcoef32) return 0;
  FPC_HIP(hipMalloc(&A->zcoef32, (size_t)kSlots * A->zpitch * sizeof(float)));
  FPC_HIP(hipMalloc(&A->zdinv32, (size_t)A->zpitch * sizeof(float)));
  const long long n = (long long)kSlots * A->zpitch;
  k_to_fp32<<<blocks_for(n), 256, 0, stream()>>>(A->zcoef, A->zcoef32, n);
  FPC_LAUNCH_CHECK();
  k_dinv<float><<<blocks_for(A->zpitch), 256, 0, stream()>>>(A->zcoef32, A->zdinv32, A->zpitch);
  FPC_LAUNCH_CHECK();
  return 0;
}

// Plans the hot copy of a z-extruded operator on a host copy of its planes
// (extrude.hpp) and uploads it; 1/diag is derived on the device as dinv is.
int make_extruded(fpc_StructMatrix_s *A) {
  const int nx = A->grid.n[0], ny = A->grid.n[1], nz = A->grid.n[2];
  const long long N = A->grid.N;
  if (!z_extrude_enabled() || nz < 5 || N <= 0) return 0;
  std::vector<double> h((size_t)kSlots * (size_t)A->pitch);
  FPC_HIP(hipMemcpyAsync(h.data(), A->coef, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream()));
  FPC_HIP(hipStreamSynchronize(stream()));
  std::vector<const double *> slots;
  for (int q = 0; q < kSlots; ++q) slots.push_back(h.data() + (size_t)q * (size_t)A->pitch);
  ExtrudePlan plan = plan_extrusion(nx, ny, nz, slots);
  if (!plan.m) return 0;
  A->zpitch = pad256(plan.xoff + N, (int)sizeof(double));
  std::vector<double> hot = hot_planes(h.data(), A->pitch, kSlots, nx, ny, nz, plan, A->zpitch);
  FPC_HIP(hipMalloc(&A->zcoef, hot.size() * sizeof(double)));
  FPC_HIP(hipMalloc(&A->zdinv, (size_t)A->zpitch * sizeof(double)));
  const std::vector<unsigned long long> words = flag_words(plan, nx, ny, nz);
  FPC_HIP(hipMalloc(&A->zxf, words.size() * sizeof(unsigned long long)));
  FPC_HIP(hipMemcpyAsync(A->zcoef, hot.data(), hot.size() * sizeof(double), hipMemcpyHostToDevice, stream()));
  FPC_HIP(hipMemcpyAsync(A->zxf, words.data(), words.size() * sizeof(unsigned long long), hipMemcpyHostToDevice,
                         stream()));
  k_dinv<double><<<blocks_for(A->zpitch), 256, 0, stream()>>>(A->zcoef, A->zdinv, A->zpitch);
  FPC_LAUNCH_CHECK();
  FPC_HIP(hipStreamSynchronize(stream()));  // hot and words are released on return
  A->zexc = plan.nexc;
  A->zmap.pl = (long long)nx * ny;
  A->zmap.nz = nz;
  A->zmap.m = plan.m;
  A->zmap.xoff = plan.xoff;
  A->zmap.xb = plan.nexc ? A->zxf : nullptr;
  return pack_extruded32(A);
}

}  // namespace

int ensure_dinv(fpc_StructMatrix_s *A) {
  if (A->dinv) return 0;
  FPC_HIP(hipMalloc(&A->dinv, (size_t)A->pitch * sizeof(double)));
  k_dinv<double><<<blocks_for(A->grid.N), 256, 0, stream()>>>(A->coef, A->dinv, A->grid.N);
  FPC_LAUNCH_CHECK();
  return make_extruded(A);
}

int ensure_fp32(fpc_StructMatrix_s *A) {
  if (A->coef32) return 0;
  FPC_HIP(hipMalloc(&A->coef32, (size_t)kSlots * A->pitch * sizeof(float)));
  FPC_HIP(hipMalloc(&A->dinv32, (size_t)A->pitch * sizeof(float)));
  long long n = (long long)kSlots * A->pitch;
  k_to_fp32<<<blocks_for(n), 256, 0, stream()>>>(A->coef, A->coef32, n);
  FPC_LAUNCH_CHECK();
  k_dinv<float><<<blocks_for(A->grid.N), 256, 0, stream()>>>(A->coef32, A->dinv32, A->grid.N);
  FPC_LAUNCH_CHECK();
  return pack_extruded32(A);
}

static void drop_derived(fpc_StructMatrix_s *A) {
  for (void *p : {(void *)A->dinv, (void *)A->coef32, (void *)A->dinv32, (void *)A->zcoef, (void *)A->zdinv,
                  (void *)A->zcoef32, (void *)A->zdinv32, (void *)A->zxf})
    if (p) (void)hipFree(p);
  A->dinv = nullptr;
  A->coef32 = nullptr;
  A->dinv32 = nullptr;
  A->zcoef = A->zdinv = nullptr;
  A->zcoef32 = A->zdinv32 = nullptr;
  A->zxf = nullptr;
  A->zpitch = A->zexc = 0;
  A->zmap = ZMap();
}

}  // namespace fpc

using namespace fpc;

// Maps an offset to its canonical slot (C,-x,+x,-y,+y,-z,+z) or -1.
static int slot_of_offset(const int *o) {
  static const int off[kSlots][3] = {{0, 0, 0}, {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
  for (int s = 0; s < kSlots; ++s)
    if (o[0] == off[s][0] && o[1] == off[s][1] && o[2] == off[s][2]) return s;
  return -1;
}

// Sub-box [il, iu] of a grid -> offsets and extents; false when outside.
static bool sub_box(const fpc_StructGrid_s &g, const int *il, const int *iu, int *o, int *e) {
  for (int d = 0; d < 3; ++d) {
    o[d] = il[d] - g.il[d];
    e[d] = iu[d] - il[d] + 1;
    if (o[d] < 0 || e[d] < 0 || o[d] + e[d] > g.n[d]) return false;
  }
  return true;
}

extern "C" {

int fpc_Initialize(void) {
  ensure_init();
  return error_bits() & FPC_ERROR_GENERIC;
}

int fpc_Finalize(void) {
  release_thread_state();
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_stream && g_own_stream) (void)hipStreamDestroy(g_stream);
  g_stream = nullptr;
  g_own_stream = false;
  g_init.store(false, std::memory_order_release);
  return 0;
}

int fpc_SetMemoryLocationDevice(void) { return 0; }
int fpc_SetExecutionPolicyDevice(void) { return 0; }
int fpc_GetError(void) { return t_state.error; }
int fpc_ClearError(int bits) {
  t_state.error &= ~bits;
  return 0;
}
const char *fpc_GetErrorMessage(void) { return t_state.msg; }

int fpc_SetStream(void *s) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_stream && g_own_stream) (void)hipStreamDestroy(g_stream);
  g_stream = (hipStream_t)s;
  g_own_stream = false;
  // NULL: a fresh library stream is created on next use
  g_init.store(s != nullptr, std::memory_order_release);
  return 0;
}

void *fpc_GetStream(void) { return (void *)stream(); }

// ---------------------------------------------------------------- grid
int fpc_StructGridCreate(int ndim, fpc_StructGrid *grid) {
  FPC_CHECK_ARG(grid, "null handle pointer");
  FPC_CHECK_ARG(ndim >= 1 && ndim <= 3, "ndim must be 1..3");
  ensure_init();
  fpc_StructGrid g = new fpc_StructGrid_s();
  g->ndim = ndim;
  *grid = g;
  return 0;
}

int fpc_StructGridSetExtents(fpc_StructGrid g, const int *il, const int *iu) {
  FPC_CHECK_ARG(g && il && iu, "null argument");
  FPC_CHECK_ARG(!g->extents, "only one box per grid is supported");
  for (int d = 0; d < 3; ++d) {
    g->il[d] = d < g->ndim ? il[d] : 0;
    g->iu[d] = d < g->ndim ? iu[d] : 0;
    FPC_CHECK_ARG(g->iu[d] >= g->il[d], "empty box");
    g->n[d] = g->iu[d] - g->il[d] + 1;
  }
  g->N = (long long)g->n[0] * g->n[1] * g->n[2];
  g->extents = true;
  return 0;
}

int fpc_StructGridAssemble(fpc_StructGrid g) {
  FPC_CHECK_ARG(g && g->extents, "grid has no extents");
  g->assembled = true;
  return 0;
}

int fpc_StructGridDestroy(fpc_StructGrid g) {
  delete g;
  return 0;
}

// ---------------------------------------------------------------- stencil
int fpc_StructStencilCreate(int ndim, int size, fpc_StructStencil *st) {
  FPC_CHECK_ARG(st, "null handle pointer");
  FPC_CHECK_ARG(size >= 1 && size <= kSlots, "stencil size must be 1..7 (7-point stencils)");
  fpc_StructStencil s = new fpc_StructStencil_s();
  s->ndim = ndim;
  s->size = size;
  *st = s;
  return 0;
}

int fpc_StructStencilSetElement(fpc_StructStencil st, int entry, const int *offset) {
  FPC_CHECK_ARG(st && offset, "null argument");
  FPC_CHECK_ARG(entry >= 0 && entry < st->size, "entry out of range");
  int o[3] = {0, 0, 0};
  for (int d = 0; d < st->ndim && d < 3; ++d) o[d] = offset[d];
  FPC_CHECK_ARG(slot_of_offset(o) >= 0, "only the 7-point offsets are supported");
  for (int d = 0; d < 3; ++d) st->off[entry][d] = o[d];
  st->set_mask |= 1 << entry;
  return 0;
}

int fpc_StructStencilDestroy(fpc_StructStencil st) {
  delete st;
  return 0;
}

// ---------------------------------------------------------------- matrix
int fpc_StructMatrixCreate(fpc_StructGrid grid, fpc_StructStencil st, fpc_StructMatrix *A) {
  FPC_CHECK_ARG(grid && st && A, "null argument");
  FPC_CHECK_ARG(grid->assembled, "grid not assembled");
  FPC_CHECK_ARG(st->set_mask == (1 << st->size) - 1, "stencil elements not all set");
  fpc_StructMatrix M = new fpc_StructMatrix_s();
  M->grid = *grid;
  M->nentries = st->size;
  for (int e = 0; e < st->size; ++e) M->slot_of_entry[e] = slot_of_offset(st->off[e]);
  *A = M;
  return 0;
}

int fpc_StructMatrixInitialize(fpc_StructMatrix A) {
  FPC_CHECK_ARG(A, "null matrix");
  if (!A->coef) {
    A->pitch = pad256(A->grid.N, (int)sizeof(double));
    FPC_HIP(hipMalloc(&A->coef, (size_t)kSlots * A->pitch * sizeof(double)));
    FPC_HIP(hipMemsetAsync(A->coef, 0, (size_t)kSlots * A->pitch * sizeof(double), stream()));
  }
  A->initialized = true;
  return 0;
}

static int matrix_box_io(fpc_StructMatrix A, const int *il, const int *iu, int nent, const int *entries,
                         const double *vin, double *vout) {
  FPC_CHECK_ARG(A && il && iu && entries, "null argument");
  FPC_CHECK_ARG(A->initialized, "matrix not initialized");
  FPC_CHECK_ARG(nent >= 1 && nent <= A->nentries, "bad number of entries");
  int o[3], e[3];
  FPC_CHECK_ARG(sub_box(A->grid, il, iu, o, e), "box outside the grid");
  int slots[kSlots];
  for (int t = 0; t < nent; ++t) {
    FPC_CHECK_ARG(entries[t] >= 0 && entries[t] < A->nentries, "entry index out of range");
    slots[t] = A->slot_of_entry[entries[t]];
  }
  long long npts = (long long)e[0] * e[1] * e[2];
  if (npts == 0) return 0;
  size_t nval = (size_t)npts * nent;
  double *stage = scratch_doubles(nval, 0);
  double *dslots = scratch_doubles(kSlots, 1);
  if (!stage || !dslots) return FPC_ERROR_MEMORY;
  int *islots = (int *)dslots;
  FPC_HIP(hipMemcpyAsync(islots, slots, sizeof(int) * nent, hipMemcpyHostToDevice, stream()));
  if (vin) FPC_HIP(hipMemcpyAsync(stage, vin, nval * sizeof(double), hipMemcpyDefault, stream()));
  k_set_matrix_box<<<blocks_for(npts), 256, 0, stream()>>>(A->coef, A->pitch, A->grid.n[0], A->grid.n[1], o[0], o[1],
                                                             o[2], e[0], e[1], e[2], nent, islots, stage, vin ? 0 : 1,
                                                             stage);
  FPC_LAUNCH_CHECK();
  if (vout) FPC_HIP(hipMemcpyAsync(vout, stage, nval * sizeof(double), hipMemcpyDefault, stream()));
  FPC_HIP(hipStreamSynchronize(stream()));
  if (vin) {
    drop_derived(A);
    A->version++;
  }
  return 0;
}

int fpc_StructMatrixSetBoxValues(fpc_StructMatrix A, const int *il, const int *iu, int nent, const int *entries,
                                 const double *values) {
  FPC_CHECK_ARG(values, "null values");
  return matrix_box_io(A, il, iu, nent, entries, values, nullptr);
}

int fpc_StructMatrixGetBoxValues(fpc_StructMatrix A, const int *il, const int *iu, int nent, const int *entries,
                                 double *values) {
  FPC_CHECK_ARG(values, "null values");
  return matrix_box_io(A, il, iu, nent, entries, nullptr, values);
}

int fpc_StructMatrixAssemble(fpc_StructMatrix A) {
  FPC_CHECK_ARG(A && A->initialized, "matrix not initialized");
  int rc = ensure_dinv(A);
  if (rc) return rc;
  FPC_HIP(hipStreamSynchronize(stream()));
  A->assembled = true;
  return 0;
}

int fpc_StructMatrixMatvec(double alpha, fpc_StructMatrix A, fpc_StructVector x, double beta, fpc_StructVector y) {
  FPC_CHECK_ARG(A && x && y, "null argument");
  FPC_CHECK_ARG(A->initialized && x->data && y->data, "objects not initialized");
  FPC_CHECK_ARG(x->grid.N == A->grid.N && y->grid.N == A->grid.N, "size mismatch");
  FPC_CHECK_ARG(x->data != y->data, "x and y must differ");
  int rc = matvec(alpha, A, x->data, beta, y->data);
  if (rc) return rc;
  FPC_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int fpc_StructMatrixGetExtruded(fpc_StructMatrix A, int *wall_planes, long long *flagged) {
  FPC_CHECK_ARG(A && wall_planes && flagged, "null argument");
  *wall_planes = A->zmap.pl ? A->zmap.m : 0;
  *flagged = A->zexc;
  return 0;
}

int fpc_StructMatrixDestroy(fpc_StructMatrix A) {
  if (!A) return 0;
  drop_derived(A);
  if (A->coef) (void)hipFree(A->coef);
  delete A;
  return 0;
}

// ---------------------------------------------------------------- vector
int fpc_StructVectorCreate(fpc_StructGrid grid, fpc_StructVector *x) {
  FPC_CHECK_ARG(grid && x, "null argument");
  FPC_CHECK_ARG(grid->assembled, "grid not assembled");
  fpc_StructVector v = new fpc_StructVector_s();
  v->grid = *grid;
  *x = v;
  return 0;
}

int fpc_StructVectorInitialize(fpc_StructVector x) {
  FPC_CHECK_ARG(x, "null vector");
  if (!x->data) {
    size_t n = (size_t)pad256(std::max(1LL, x->grid.N), (int)sizeof(double));
    FPC_HIP(hipMalloc(&x->data, n * sizeof(double)));
    FPC_HIP(hipMemsetAsync(x->data, 0, n * sizeof(double), stream()));
    FPC_HIP(hipStreamSynchronize(stream()));
  }
  x->initialized = true;
  return 0;
}

static int vector_box_io(fpc_StructVector x, const int *il, const int *iu, const double *vin, double *vout) {
  FPC_CHECK_ARG(x && il && iu, "null argument");
  FPC_CHECK_ARG(x->initialized, "vector not initialized");
  int o[3], e[3];
  FPC_CHECK_ARG(sub_box(x->grid, il, iu, o, e), "box outside the grid");
  long long npts = (long long)e[0] * e[1] * e[2];
  if (npts == 0) return 0;
  bool whole = (npts == x->grid.N);
  if (whole) {
    if (vin) FPC_HIP(hipMemcpyAsync(x->data, vin, (size_t)npts * sizeof(double), hipMemcpyDefault, stream()));
    else FPC_HIP(hipMemcpyAsync(vout, x->data, (size_t)npts * sizeof(double), hipMemcpyDefault, stream()));
  } else {
    double *stage = scratch_doubles((size_t)npts, 2);
    if (!stage) return FPC_ERROR_MEMORY;
    if (vin) FPC_HIP(hipMemcpyAsync(stage, vin, (size_t)npts * sizeof(double), hipMemcpyDefault, stream()));
    k_vector_box<<<blocks_for(npts), 256, 0, stream()>>>(x->data, x->grid.n[0], x->grid.n[1], o[0], o[1], o[2], e[0],
                                                        e[1], e[2], vin ? stage : nullptr, vin ? nullptr : stage);
    FPC_LAUNCH_CHECK();
    if (vout) FPC_HIP(hipMemcpyAsync(vout, stage, (size_t)npts * sizeof(double), hipMemcpyDefault, stream()));
  }
  FPC_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int fpc_StructVectorSetBoxValues(fpc_StructVector x, const int *il, const int *iu, const double *values) {
  FPC_CHECK_ARG(values, "null values");
  return vector_box_io(x, il, iu, values, nullptr);
}

int fpc_StructVectorGetBoxValues(fpc_StructVector x, const int *il, const int *iu, double *values) {
  FPC_CHECK_ARG(values, "null values");
  return vector_box_io(x, il, iu, nullptr, values);
}

int fpc_StructVectorSetConstantValues(fpc_StructVector x, double value) {
  FPC_CHECK_ARG(x && x->initialized, "vector not initialized");
  if (x->grid.N > 0) {
    k_fill<<<blocks_for(x->grid.N), 256, 0, stream()>>>(x->data, x->grid.N, value);
    FPC_LAUNCH_CHECK();
  }
  FPC_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int fpc_StructVectorAssemble(fpc_StructVector x) {
  FPC_CHECK_ARG(x && x->initialized, "vector not initialized");
  return 0;
}

int fpc_StructVectorGetDataPointer(fpc_StructVector x, double **p) {
  FPC_CHECK_ARG(x && p && x->initialized, "vector not initialized");
  *p = x->data;
  return 0;
}

int fpc_StructVectorDestroy(fpc_StructVector x) {
  if (!x) return 0;
  if (x->data) (void)hipFree(x->data);
  delete x;
  return 0;
}

// ---------------------------------------------------------------- helpers (hypre_struct_utils.hpp:1-26)
fpc_StructGrid fpc_CreateGrid3D(int nx, int ny, int nz) {
  fpc_StructGrid g = nullptr;
  int il[3] = {1, 1, 1}, iu[3] = {nx, ny, nz};
  if (fpc_StructGridCreate(3, &g)) return nullptr;
  if (fpc_StructGridSetExtents(g, il, iu) || fpc_StructGridAssemble(g)) {
    fpc_StructGridDestroy(g);
    return nullptr;
  }
  return g;
}

fpc_StructStencil fpc_CreateStencil3D(void) {
  fpc_StructStencil st = nullptr;
  int offs[7][3] = {{0, 0, 0}, {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
  if (fpc_StructStencilCreate(3, 7, &st)) return nullptr;
  for (int s = 0; s < 7; ++s)
    if (fpc_StructStencilSetElement(st, s, offs[s])) {
      fpc_StructStencilDestroy(st);
      return nullptr;
    }
  return st;
}

int fpc_SetMatrix3D(fpc_StructMatrix A, int nx, int ny, int nz, const double *vals) {
  int il[3] = {1, 1, 1}, iu[3] = {nx, ny, nz}, entries[7] = {0, 1, 2, 3, 4, 5, 6};
  int rc = fpc_StructMatrixInitialize(A);
  if (!rc) rc = fpc_StructMatrixSetBoxValues(A, il, iu, 7, entries, vals);
  if (!rc) rc = fpc_StructMatrixAssemble(A);
  return rc;
}

fpc_StructMatrix fpc_CreateMatrix3D(fpc_StructGrid grid, fpc_StructStencil st, int nx, int ny, int nz,
                                    const double *vals) {
  fpc_StructMatrix A = nullptr;
  if (fpc_StructMatrixCreate(grid, st, &A)) return nullptr;
  if (fpc_SetMatrix3D(A, nx, ny, nz, vals)) {
    fpc_StructMatrixDestroy(A);
    return nullptr;
  }
  return A;
}

int fpc_SetVector3D(fpc_StructVector x, int nx, int ny, int nz, const double *vals) {
  int il[3] = {1, 1, 1}, iu[3] = {nx, ny, nz};
  int rc = fpc_StructVectorSetBoxValues(x, il, iu, vals);
  if (!rc) rc = fpc_StructVectorAssemble(x);
  return rc;
}

int fpc_GetVector3D(fpc_StructVector x, int nx, int ny, int nz, double *vals) {
  int il[3] = {1, 1, 1}, iu[3] = {nx, ny, nz};
  return fpc_StructVectorGetBoxValues(x, il, iu, vals);
}

fpc_StructVector fpc_CreateVector3D(fpc_StructGrid grid) {
  fpc_StructVector x = nullptr;
  if (fpc_StructVectorCreate(grid, &x)) return nullptr;
  if (fpc_StructVectorInitialize(x)) {
    fpc_StructVectorDestroy(x);
    return nullptr;
  }
  return x;
}

// ---------------------------------------------------------------- memory helpers (hypre_struct_utils.hpp:27-43)
static int dmalloc(void **p, size_t bytes) {
  FPC_CHECK_ARG(p, "null pointer");
  ensure_init();
  *p = nullptr;
  FPC_HIP(hipMalloc(p, bytes ? bytes : 1));
  return 0;
}

int fpc_MallocManagedC(double **p, size_t n) { return dmalloc((void **)p, n * sizeof(double)); }
int fpc_MallocManagedD(double **p, size_t n) { return dmalloc((void **)p, n * sizeof(double)); }
int fpc_MallocManagedI(int **p, size_t n) { return dmalloc((void **)p, n * sizeof(int)); }
int fpc_MallocManagedB(unsigned char **p, size_t n) { return dmalloc((void **)p, n); }

static int dcopyvec(void **p, const void *src, size_t bytes) {
  int rc = dmalloc(p, bytes);
  if (rc) return rc;
  if (bytes) {
    FPC_HIP(hipMemcpyAsync(*p, src, bytes, hipMemcpyDefault, stream()));
    FPC_HIP(hipStreamSynchronize(stream()));
  }
  return 0;
}

int fpc_CopyVecToManagedC(double **p, const double *s, size_t n) { return dcopyvec((void **)p, s, n * sizeof(double)); }
int fpc_CopyVecToManagedD(double **p, const double *s, size_t n) { return dcopyvec((void **)p, s, n * sizeof(double)); }
int fpc_CopyVecToManagedI(int **p, const int *s, size_t n) { return dcopyvec((void **)p, s, n * sizeof(int)); }
int fpc_CopyVecToManagedB(unsigned char **p, const unsigned char *s, size_t n) { return dcopyvec((void **)p, s, n); }

int fpc_Free(void *p) {
  if (p) FPC_HIP(hipFree(p));
  return 0;
}

int fpc_MemcpyToHost(void *dst, const void *src, size_t bytes) {
  if (!bytes) return 0;
  FPC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, stream()));
  FPC_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int fpc_MemcpyToDevice(void *dst, const void *src, size_t bytes) {
  if (!bytes) return 0;
  FPC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, stream()));
  FPC_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int fpc_SetTuning(const char *key, long long value) {
  FPC_CHECK_ARG(key, "null key");
  std::string k(key);
  TileTuning &t = tile_tuning();
  if (k == "tile_min_points") {
    FPC_CHECK_ARG(value >= 0, "tile_min_points must be >= 0");
    t.min_points = value;
  } else if (k == "stencil_kz") {
    FPC_CHECK_ARG(value >= 0 && value <= 1024, "stencil_kz must be in [0, 1024]");
    t.kz = (int)value;
  } else if (k == "z_extrude") {
    FPC_CHECK_ARG(value == 0 || value == 1, "z_extrude must be 0 or 1");
    t.z_extrude = (int)value;
  } else if (k == "mg_wide_finest") {
    FPC_CHECK_ARG(value == 0 || value == 1, "mg_wide_finest must be 0 or 1");
    t.mg_wide_finest = (int)value;
  } else if (k == "p_flexible") {
    FPC_CHECK_ARG(value == 0 || value == 1, "p_flexible must be 0 or 1");
    t.p_flexible = (int)value;
  } else if (k == "mg_graph") {
    FPC_CHECK_ARG(value == 0 || value == 1, "mg_graph must be 0 or 1");
    t.mg_graph = (int)value;
  } else if (k == "mg_block") {
    FPC_CHECK_ARG(value == 64 || value == 128 || value == 256, "mg_block must be 64, 128 or 256");
    t.mg_block = (int)value;
  } else {
    FPC_CHECK_ARG(false, "unknown tuning key");
  }
  return 0;
}

int fpc_ProfileSample(int every) {
  FPC_CHECK_ARG(every >= 1, "every must be >= 1");
  g_prof_every = every;
  g_prof_seen = 0;
  return 0;
}

int fpc_ProfileEnable(const char *names) {
  g_prof_filter = names ? names : "";
  g_prof_cls.clear();
  g_prof = !g_prof_filter.empty();
  return 0;
}

int fpc_ProfileReset(void) {
  // The records go and so do their events: recorded events left alive slow the
  // runtime's later launches and syncs (measured at 512x96x96: 1,700 of them cost
  // 0.7% of every later step), so a timed region starts with none.
  if (!g_prof_pool.empty()) (void)hipStreamSynchronize(stream());
  for (auto &pr : g_prof_pool) {
    (void)hipEventDestroy(pr.first);
    (void)hipEventDestroy(pr.second);
  }
  g_prof_pool.clear();
  g_prof_recs.clear();
  g_prof_used = 0;
  g_prof_open = nullptr;
  return 0;
}

int fpc_ProfileReport(char *buf, int buflen) {
  FPC_HIP(hipStreamSynchronize(stream()));
  std::vector<std::string> names;
  std::vector<long long> cnt;
  std::vector<double> ms, bytes;
  for (auto &r : g_prof_recs) {
    if (r.dropped) continue;
    float t = 0.0f;
    if (hipEventElapsedTime(&t, r.e0, r.e1) != hipSuccess) t = 0.0f;
    size_t q = 0;
    while (q < names.size() && names[q] != r.cls) ++q;
    if (q == names.size()) {
      names.push_back(r.cls);
      cnt.push_back(0);
      ms.push_back(0.0);
      bytes.push_back(0.0);
    }
    cnt[q] += 1;
    ms[q] += t;
    bytes[q] += r.bytes;
  }
  std::string out;
  char line[256];
  for (size_t q = 0; q < names.size(); ++q) {
    snprintf(line, sizeof(line), "%s %lld %.6f %.6e\n", names[q].c_str(), cnt[q], ms[q], bytes[q]);
    out += line;
  }
  if (buf && buflen > 0) {
    size_t n = std::min(out.size(), (size_t)buflen - 1);
    std::memcpy(buf, out.data(), n);
    buf[n] = 0;
  }
  return (int)out.size() + 1;
}

int fpc_DeviceSynchronize(void) {
  FPC_HIP(hipStreamSynchronize(stream()));
  return 0;
}

}  // extern "C"

// batched.hip -- many independent Struct GMRES solves advanced by shared
// launches (libfpc).  No counterpart in the reference, which solves one
// channel per process.
//
// A group holds the set-up solvers of many systems (the u, v, w momentum
// solves of every channel of an ensemble, or their pressure solves).  Every
// kernel of an Arnoldi step, and every level of the PFMG-style V-cycle, is one
// launch over all systems of the group:
//  * pointwise kernels (matvec, MG down/restrict/up, coarsest GEMV) run over a
//    packed block table: block b serves system item_of_block[b], points
//    (b - first_block[item]) * 256 + thread; systems of different sizes share
//    the launch with no idle blocks;
//  * reductions run on a (blocks, systems) grid with each system's own
//    red_blocks(N) partition and the same per-block bodies as the solver
//    alone (krylov_device.hpp, mg_device.hpp), so sums are formed in the same
//    order;
//  * systems whose device state says "done" return at once, as in the
//    speculative Arnoldi batches of one solver.
// The host runs fpc_StructGMRESSolve's restart and convergence decisions for
// every system in lockstep: a step that one system needs and another does not
// is skipped by the latter's device flag.  Each system therefore computes bit
// for bit what its own solve computes (same iteration counts, same iterates).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "batched.hpp"
#include "fpc_internal.hpp"
#include "krylov_device.hpp"
#include "mg_device.hpp"
#include "stencil_tile.hpp"

namespace fpc {
namespace {

using kry::KM;
using kry::NB;
using kry::State;
using kry::Status;

constexpr int kBMax = 768;  // systems per group

struct BSys {
  const double *c;  // the operator's planes as the z-marching kernels read them (rows by zm)
  long long pitch;
  ZMap zm;
  const double *fc;  // the full planes (one-point kernels)
  long long fpitch;
  int nx, ny, nz, nred;
  long long N, vp;
  const double *dinv;       // fused Jacobi, stored like c, or null
  const double *dinv_full;  // the same 1/diag at every point (one-point matvec, solution update), or null
  double *V, *Z, *W, *part;
  State *S;
  const double *b;
  double *x;
  int flex;  // Z holds one preconditioned vector per Arnoldi step (flexible GMRES), else one
};

// The preconditioned vector of Arnoldi step `step` (flexible: its own slot).
__device__ inline double *z_slot(const BSys &s, int step) { return s.flex ? s.Z + (long long)step * s.vp : s.Z; }

struct BLev {
  const double *c, *dinv, *wm, *wp;  // z-marching kernels: rows by zm
  long long pitch;
  ZMap zm;
  const double *fc, *fdinv, *fwm, *fwp;  // one-point kernels: the full arrays
  long long fpitch;
  double *b, *r, *y;  // level >= 1 (level 0 reads the call's input, writes Z)
  double *bc;         // coarse b (restriction target)
  const double *yc;   // coarse y (prolongation source)
  int nx, ny, nz, cdir;
  long long N, Nc;
  double w;
  int cnx, cny, cnz, ncd;
  const double *inv;  // coarsest level: FP64 inverse
  int ncoarse;
  int sys, lvl;
};

// A subset of the group's systems, passed by value.
struct Sub {
  int n;
  unsigned short idx[kBMax];
};

// Packed block table: block b serves item[b]; first[item] is the item's first block.
struct Table {
  int *item = nullptr, *first = nullptr;
  int nblocks = 0;
};

// One workgroup of a z-marching launch: tile (bx, by) = (bxy & 0xffff, bxy >> 16)
// of item, planes k0 .. k1-1.
struct TileEnt {
  int item, bxy, k0, k1;
};

struct TileTable {
  TileEnt *ent = nullptr;
  int nblocks = 0;
};

// Tile configurations tried by the batch autotuner (BatchGroup::autotune): boxes
// of at least tile_min points (and nx >= 32, ny >= 8) run the z-marching kernels
// in z-chunks of at most kz planes, the others the one-point kernels.  With many
// systems per launch the tiles do not lack workgroups, unlike one small box
// alone.  Every configuration computes the same bits.
struct TileConfig {
  int kz;
  long long tile_min;
};
constexpr TileConfig kTileConfigs[] = {{4, 16384}, {8, 16384}, {16, 16384}, {2, 16384}, {4, 65536}, {4, 1LL << 40}};
constexpr int kTY = 8;

__device__ inline void ijk_of(long long t, int nx, int ny, int &i, int &j, int &k) {
  i = (int)(t % nx);
  const long long q = t / nx;
  j = (int)(q % ny);
  k = (int)(q / ny);
}

// ------------------------------------------------------------ Krylov kernels
// zero: every system starts from x = 0, so r = b (copied, the same partials bit for bit)
__global__ void __launch_bounds__(256) k_b_residual(const BSys *sys, Sub sub, int out_mode, int zero) {
  const BSys &s = sys[sub.idx[blockIdx.y]];
  if ((int)blockIdx.x >= s.nred) return;
  if (blockIdx.z == 0) {
    double *r = (out_mode == 0) ? s.V : (out_mode == 1 ? s.W : nullptr);
    if (zero && r) kry::copy_dot_body(s.b, r, s.N, s.part, blockIdx.x, s.nred);
    else kry::residual_body(s.fc, s.fpitch, FullMap(), s.nx, s.ny, s.nz, s.b, s.x, r, s.part, blockIdx.x, s.nred);
  } else {
    kry::dot_body(s.b, s.b, s.N, s.part + NB, blockIdx.x, s.nred);
  }
}

// out[q * nvals + v] = sum of the partials of value v of system sub[q]
__global__ void __launch_bounds__(1024) k_b_reduce(const BSys *sys, Sub sub, int nvals, double *out) {
  const BSys &s = sys[sub.idx[blockIdx.y]];
  kry::reduce_body(s.part + (long long)blockIdx.x * NB, s.nred, out + (long long)blockIdx.y * nvals + blockIdx.x);
}

__global__ void k_b_copy_w_v0(const BSys *sys, Sub sub) {
  const BSys &s = sys[sub.idx[blockIdx.y]];
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < s.N; t += (long long)gridDim.x * blockDim.x)
    s.V[t] = s.W[t];
}

// args per subset entry: r_norm, eps, iter, maxit, kdim
__global__ void k_b_cycle_start(const BSys *sys, Sub sub, const double *args) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= sub.n) return;
  const double *a = args + 5 * q;
  kry::cycle_start_body(sys[sub.idx[q]].S, a[0], a[1], (int)a[2], (int)a[3], (int)a[4]);
}

__global__ void k_b_set_done(const BSys *sys, int n, int v) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) sys[q].S->st.done = v;
}

__global__ void __launch_bounds__(256) k_b_matvec(const BSys *sys, const int *item, const int *first, int use_z) {
  const int q = item[blockIdx.x];
  const BSys &s = sys[q];
  State *S = s.S;
  if (S->st.done) return;
  const long long t = (long long)(blockIdx.x - first[q]) * 256 + threadIdx.x;
  if (t >= s.N) return;
  const int cur = S->st.i;
  int i, j, k;
  ijk_of(t, s.nx, s.ny, i, j, k);
  const double *z = use_z ? z_slot(s, cur) : s.V + (long long)cur * s.vp;
  const double sg = S->sigma[cur];
  const double v = kry::arnoldi_row(s.fc, s.fpitch, FullMap(), s.nx, s.ny, s.nz, z, s.dinv_full, i, j, k, t);
  s.V[(long long)(cur + 1) * s.vp + t] = sg * v;
}

__global__ void __launch_bounds__(256) k_b_multidot(const BSys *sys) {
  const BSys &s = sys[blockIdx.y];
  if (s.S->st.done || (int)blockIdx.x >= s.nred) return;
  kry::multidot_body(s.V, s.vp, s.N, s.S, s.part, blockIdx.x, s.nred);
}

__global__ void __launch_bounds__(256) k_b_reduce_h(const BSys *sys) {
  const BSys &s = sys[blockIdx.y];
  kry::reduce_h_body(s.S, s.part, s.nred, blockIdx.x);
}

__global__ void __launch_bounds__(256) k_b_orth_update(const BSys *sys) {
  const BSys &s = sys[blockIdx.y];
  if (s.S->st.done || (int)blockIdx.x >= s.nred) return;
  kry::orth_update_body(s.V, s.vp, s.N, s.S, s.part, blockIdx.x, s.nred);
}

__global__ void __launch_bounds__(256) k_b_tail(const BSys *sys) {
  const BSys &s = sys[blockIdx.x];
  kry::arnoldi_tail_body(s.S, s.part, s.nred);
}

__global__ void k_b_status(const BSys *sys, int n, Status *out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) out[q] = sys[q].S->st;
}

__global__ void k_b_backsolve(const BSys *sys, Sub sub) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < sub.n) kry::backsolve_body(sys[sub.idx[q]].S);
}

// mode 0: x += [dinv] sum_j y_j V_j;  mode 1: W = sum_j y_j V_j;  mode 2: x += sum_j y_j Z_j
__global__ void __launch_bounds__(256) k_b_combine(const BSys *sys, Sub sub, int mode) {
  __shared__ double y[KM + 1];
  const BSys &s = sys[sub.idx[blockIdx.y]];
  const int i = s.S->st.i;
  for (int q = threadIdx.x; q < i; q += blockDim.x) y[q] = s.S->y[q];
  __syncthreads();
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < s.N; t += (long long)gridDim.x * blockDim.x) {
    if (mode == 0) {
      const double v = kry::combine_point(s.V, s.vp, t, y, i, s.dinv_full);
      s.x[t] = s.x[t] + v;
    } else if (mode == 2) {
      const double v = kry::combine_point<double>(s.Z, s.vp, t, y, i, nullptr);
      s.x[t] = s.x[t] + v;
    } else {
      s.W[t] = kry::combine_point<double>(s.V, s.vp, t, y, i, nullptr);
    }
  }
}

__global__ void k_b_axpy_z(const BSys *sys, Sub sub) {
  const BSys &s = sys[sub.idx[blockIdx.y]];
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < s.N; t += (long long)gridDim.x * blockDim.x)
    s.x[t] += s.Z[t];
}

__global__ void k_b_restart_coefs(const BSys *sys, Sub sub) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < sub.n) kry::restart_coefs_body(sys[sub.idx[q]].S);
}

__global__ void __launch_bounds__(256) k_b_restart_vec(const BSys *sys, Sub sub) {
  __shared__ double y[KM + 1];
  const BSys &s = sys[sub.idx[blockIdx.y]];
  const int i = s.S->st.i;
  for (int q = threadIdx.x; q <= i; q += blockDim.x) y[q] = s.S->y[q];
  __syncthreads();
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < s.N; t += (long long)gridDim.x * blockDim.x)
    s.V[t] = kry::restart_point(s.V, s.vp, t, y, i);
}

// flags[q] = 0 for the systems of sub, 1 for the others (one block)
__global__ void __launch_bounds__(1024) k_b_flags(Sub sub, int n, int *flags) {
  for (int q = threadIdx.x; q < n; q += blockDim.x) flags[q] = 1;
  __syncthreads();
  for (int e = threadIdx.x; e < sub.n; e += blockDim.x) flags[sub.idx[e]] = 0;
}

// ------------------------------------------------------------ V-cycle kernels
// Level 0 reads V[step] (in_mode 0) or W (in_mode 1) and writes Z; a system
// skips when flags[sys] != 0, or, without flags, when its solve state is done.
__device__ inline bool lev_skip(const BLev &L, const BSys &s, const int *flags) {
  return flags ? flags[L.sys] != 0 : s.S->st.done != 0;
}
__device__ inline const double *lev_b(const BLev &L, const BSys &s, int in_mode, int step) {
  return L.lvl == 0 ? (in_mode == 0 ? s.V + (long long)step * s.vp : s.W) : L.b;
}

// one-point levels: residual and restriction in one launch, one thread per coarse point
__global__ void __launch_bounds__(256) k_b_down_restrict(const BLev *lev, const BSys *sys, const int *item,
                                                         const int *first, int in_mode, int step, const int *flags) {
  const int q = item[blockIdx.x];
  const BLev &L = lev[q];
  const BSys &s = sys[L.sys];
  if (lev_skip(L, s, flags)) return;
  const long long t = (long long)(blockIdx.x - first[q]) * 256 + threadIdx.x;
  if (t >= L.Nc) return;
  int i, j, k;
  ijk_of(t, L.cnx, L.cny, i, j, k);
  L.bc[t] = mgd::down_restrict_point<double>(L.fc, L.fpitch, FullMap(), L.nx, L.ny, L.nz, L.fdinv,
                                             lev_b(L, s, in_mode, step), L.w, L.fwm, L.fwp, L.cdir, i, j, k);
}

__global__ void __launch_bounds__(256) k_b_restrict(const BLev *lev, const BSys *sys, const int *item, const int *first,
                                                    const int *flags) {
  const int q = item[blockIdx.x];
  const BLev &L = lev[q];
  if (lev_skip(L, sys[L.sys], flags)) return;
  const long long t = (long long)(blockIdx.x - first[q]) * 256 + threadIdx.x;
  if (t >= L.Nc) return;
  int i, j, k;
  ijk_of(t, L.cnx, L.cny, i, j, k);
  L.bc[t] = mgd::restrict_point<double>(L.nx, L.ny, L.nz, L.cdir, L.r, L.fwm, L.fwp, FullMap(), i, j, k);
}

// coarsest levels: four rows per block, one wavefront per row
__global__ void __launch_bounds__(256) k_b_gemv(const BLev *lev, const BSys *sys, const int *item, const int *first,
                                                int in_mode, int step, const int *flags) {
  const int q = item[blockIdx.x];
  const BLev &L = lev[q];
  const BSys &s = sys[L.sys];
  if (lev_skip(L, s, flags)) return;
  const int r = (blockIdx.x - first[q]) * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (r >= L.ncoarse) return;
  const double v = mgd::gemv_row<double>(L.inv, L.ncoarse, lev_b(L, s, in_mode, step), r, lane);
  double *y = L.lvl == 0 ? z_slot(s, step) : L.y;
  if (lane == 0) y[r] = v;
}

__global__ void __launch_bounds__(256) k_b_up(const BLev *lev, const BSys *sys, const int *item, const int *first,
                                              int in_mode, int step, const int *flags) {
  const int q = item[blockIdx.x];
  const BLev &L = lev[q];
  const BSys &s = sys[L.sys];
  if (lev_skip(L, s, flags)) return;
  const long long t = (long long)(blockIdx.x - first[q]) * 256 + threadIdx.x;
  if (t >= L.N) return;
  int i, j, k;
  ijk_of(t, L.nx, L.ny, i, j, k);
  const double *b = lev_b(L, s, in_mode, step);
  double v;
  if (L.cdir == 0)
    v = mgd::up_point<double, 0>(L.fc, L.fpitch, FullMap(), L.nx, L.ny, L.nz, L.fdinv, b, L.fwm, L.fwp, L.yc, L.cnx,
                                 L.cny, L.ncd, L.w, i, j, k, t);
  else if (L.cdir == 1)
    v = mgd::up_point<double, 1>(L.fc, L.fpitch, FullMap(), L.nx, L.ny, L.nz, L.fdinv, b, L.fwm, L.fwp, L.yc, L.cnx,
                                 L.cny, L.ncd, L.w, i, j, k, t);
  else
    v = mgd::up_point<double, 2>(L.fc, L.fpitch, FullMap(), L.nx, L.ny, L.nz, L.fdinv, b, L.fwm, L.fwp, L.yc, L.cnx,
                                 L.cny, L.ncd, L.w, i, j, k, t);
  (L.lvl == 0 ? z_slot(s, step) : L.y)[t] = v;
}

// ------------------------------------------------------------ z-marching forms
template <bool kDinv>
__global__ void __launch_bounds__(kTileX *kTY) k_bt_matvec(const BSys *sys, const TileEnt *tab, int use_z) {
  const TileEnt e = tab[blockIdx.x];
  const BSys &s = sys[e.item];
  State *S = s.S;
  if (S->st.done) return;
  const int cur = S->st.i;
  kry::ArnoldiOp<kDinv> op{s.c, s.pitch, s.nx, s.ny, use_z ? z_slot(s, cur) : s.V + (long long)cur * s.vp, s.dinv,
                           s.V + (long long)(cur + 1) * s.vp, S->sigma[cur], s.zm};
  march_tile<double, kTileX, kTY>(op, s.nx, s.ny, s.nz, e.k0, e.k1, e.bxy & 0xffff, e.bxy >> 16);
}

__global__ void __launch_bounds__(kTileX *kTY) k_bt_down(const BLev *lev, const BSys *sys, const TileEnt *tab,
                                                         int in_mode, int step, const int *flags) {
  const TileEnt e = tab[blockIdx.x];
  const BLev &L = lev[e.item];
  const BSys &s = sys[L.sys];
  if (lev_skip(L, s, flags)) return;
  const double *b = lev_b(L, s, in_mode, step);
  const int bx = e.bxy & 0xffff, by = e.bxy >> 16;
  if (L.cdir == 2) {
    // residual restricted in z on the fly (no separate restriction)
    mgd::DownRestrictZOp<double> op;
    op.c = L.c; op.pitch = L.pitch; op.nx = L.nx; op.ny = L.ny; op.dinv = L.dinv; op.b = b;
    op.w = L.w; op.wm = L.wm; op.wp = L.wp; op.bc = L.bc; op.nz = L.nz; op.ncz = L.cnz; op.k0 = e.k0;
    op.zm = L.zm;
    march_tile<double, kTileX, kTY>(op, L.nx, L.ny, L.nz, e.k0, e.k1 < L.nz ? e.k1 + 1 : e.k1, bx, by);
  } else {
    mgd::DownOp<double> op{L.c, L.pitch, L.nx, L.ny, L.dinv, b, L.r, L.w, L.zm};
    march_tile<double, kTileX, kTY>(op, L.nx, L.ny, L.nz, e.k0, e.k1, bx, by);
  }
}

__global__ void __launch_bounds__(kTileX *kTY) k_bt_up(const BLev *lev, const BSys *sys, const TileEnt *tab,
                                                       int in_mode, int step, const int *flags) {
  const TileEnt e = tab[blockIdx.x];
  const BLev &L = lev[e.item];
  const BSys &s = sys[L.sys];
  if (lev_skip(L, s, flags)) return;
  const double *b = lev_b(L, s, in_mode, step);
  double *y = L.lvl == 0 ? z_slot(s, step) : L.y;
  const int bx = e.bxy & 0xffff, by = e.bxy >> 16;
  if (L.cdir == 0) {
    mgd::UpOp<double, 0> op{L.c, L.pitch, L.nx, L.ny, L.dinv, b, L.wm, L.wp, L.yc, L.cnx, L.cny, L.ncd, y, L.w, L.zm};
    march_tile<double, kTileX, kTY>(op, L.nx, L.ny, L.nz, e.k0, e.k1, bx, by);
  } else if (L.cdir == 1) {
    mgd::UpOp<double, 1> op{L.c, L.pitch, L.nx, L.ny, L.dinv, b, L.wm, L.wp, L.yc, L.cnx, L.cny, L.ncd, y, L.w, L.zm};
    march_tile<double, kTileX, kTY>(op, L.nx, L.ny, L.nz, e.k0, e.k1, bx, by);
  } else {
    mgd::UpOp<double, 2> op{L.c, L.pitch, L.nx, L.ny, L.dinv, b, L.wm, L.wp, L.yc, L.cnx, L.cny, L.ncd, y, L.w, L.zm};
    march_tile<double, kTileX, kTY>(op, L.nx, L.ny, L.nz, e.k0, e.k1, bx, by);
  }
}

// ------------------------------------------------------------ non-finite check
__global__ void k_b_nonfinite(const double *const *arr, const long long *len, int *flags) {
  __shared__ int any;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  const double *a = arr[blockIdx.y];
  const long long n = len[blockIdx.y];
  int bad = 0;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
    bad |= !isfinite(a[t]);
  if (bad) any = 1;
  __syncthreads();
  if (threadIdx.x == 0 && any) atomicOr(&flags[blockIdx.y], 1);
}

__global__ void k_b_zero_int(int *f, int n) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) f[q] = 0;
}

template <class T>
int dupload(T **d, const std::vector<T> &h) {
  FPC_HIP(hipMalloc((void **)d, std::max<size_t>(1, h.size()) * sizeof(T)));
  if (!h.empty()) FPC_HIP(hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

// Table over items[e] (ids below nids) with count[e] points, pts_per_block points per block.
int make_table(const std::vector<int> &items, const std::vector<long long> &count, long long pts_per_block, int nids,
               Table &t) {
  std::vector<int> item, first((size_t)std::max(1, nids), 0);
  for (size_t e = 0; e < items.size(); ++e) {
    first[(size_t)items[e]] = (int)item.size();
    const long long nb = (count[e] + pts_per_block - 1) / pts_per_block;
    for (long long b = 0; b < nb; ++b) item.push_back(items[e]);
  }
  t.nblocks = (int)item.size();
  int rc = dupload(&t.item, item);
  if (!rc) rc = dupload(&t.first, first);
  return rc;
}

bool tiled_box(int nx, int ny, long long N, long long tile_min) { return nx >= 32 && ny >= 8 && N >= tile_min; }

// z-chunks of at most kzmax planes, of even length (z-restriction pairs stay in a chunk)
int tile_chunk(int nz, int kzmax) {
  const int nch = (nz + kzmax - 1) / kzmax;
  int kz = (nz + nch - 1) / nch;
  return kz + (kz & 1);
}

int make_tile_table(const std::vector<int> &items, const std::vector<int> &nx, const std::vector<int> &ny,
                    const std::vector<int> &nz, int kzmax, TileTable &t) {
  std::vector<TileEnt> ent;
  for (size_t e = 0; e < items.size(); ++e) {
    const int tx = (nx[e] + kTileX - 1) / kTileX, ty = (ny[e] + kTY - 1) / kTY, kz = tile_chunk(nz[e], kzmax);
    for (int k0 = 0; k0 < nz[e]; k0 += kz)
      for (int by = 0; by < ty; ++by)
        for (int bx = 0; bx < tx; ++bx) ent.push_back(TileEnt{items[e], bx | (by << 16), k0, std::min(k0 + kz, nz[e])});
  }
  t.nblocks = (int)ent.size();
  return dupload(&t.ent, ent);
}

void free_tiles(TileTable &t) {
  if (t.ent) (void)hipFree(t.ent);
  t = TileTable{};
}

void free_table(Table &t) {
  if (t.item) (void)hipFree(t.item);
  if (t.first) (void)hipFree(t.first);
  t = Table{};
}

unsigned grid_blocks(long long n) { return (unsigned)std::max(1LL, std::min((n + 255) / 256, 256LL)); }

}  // namespace

// ------------------------------------------------------------ group
class BatchGroup {
 public:
  int n = 0;
  int pc = 0;
  int flex = 0;  // flexible GMRES (PFMG-style preconditioner): stored preconditioned vectors
  std::vector<GmresBatchView> view;
  BSys *dsys = nullptr;
  Table pts;          // one-point matvec over the small systems
  TileTable ptsT;     // z-marching matvec over the others
  // V-cycle: per level l, the items (fine levels l of systems deeper than l+1)
  std::vector<BLev> hlev;
  BLev *dlev = nullptr;
  std::vector<Table> down, dres, restr;  // one-point up (fine points), one-point down + restriction
                                         // (coarse points), restriction of tiled x/y-coarsened levels
  std::vector<TileTable> downT;    // z-marching down/up
  Table coarse;
  int *flags = nullptr;
  // host staging (pinned)
  Status *hst = nullptr;
  Status *dst = nullptr;
  double *hscal = nullptr, *dscal = nullptr;
  double *hargs = nullptr, *dargs = nullptr;
  long long maxN = 0;
  int maxnred = 1;
  std::vector<BSys> hsys;
  std::vector<int> nlev, base;  // levels per system, first level index
  TileConfig tiles{4, 16384};
  // V-cycles of the Arnoldi steps captured as graphs, by step ("mg_graph" tuning key);
  // freed whenever the launch tables they hold are rebuilt
  std::vector<std::pair<int, hipGraphExec_t>> vgraphs;

  void free_graphs() {
    for (auto &g : vgraphs) (void)hipGraphExecDestroy(g.second);
    vgraphs.clear();
  }

  ~BatchGroup() {
    free_graphs();
    if (dsys) (void)hipFree(dsys);
    if (dlev) (void)hipFree(dlev);
    free_table(pts);
    free_tiles(ptsT);
    for (auto &t : down) free_table(t);
    for (auto &t : dres) free_table(t);
    for (auto &t : downT) free_tiles(t);
    for (auto &t : restr) free_table(t);
    free_table(coarse);
    if (flags) (void)hipFree(flags);
    if (dst) (void)hipFree(dst);
    if (dscal) (void)hipFree(dscal);
    if (dargs) (void)hipFree(dargs);
    if (hst) (void)hipHostFree(hst);
    if (hscal) (void)hipHostFree(hscal);
    if (hargs) (void)hipHostFree(hargs);
  }

  Sub sub_of(const std::vector<int> &ids) const {
    Sub s;
    s.n = (int)ids.size();
    for (int q = 0; q < s.n; ++q) s.idx[q] = (unsigned short)ids[(size_t)q];
    return s;
  }

  // ||b - A x||^2 (r into V0 when out_mode 0, W when 1) and optionally ||b||^2 of the
  // systems ids; values in hscal[q * 2 + v]
  int true_residual(const std::vector<int> &ids, int out_mode, bool with_b, bool zero = false) {
    if (ids.empty()) return 0;
    const Sub s = sub_of(ids);
    const int nv = with_b ? 2 : 1;
    k_b_residual<<<dim3(maxnred, s.n, nv), 256, 0, stream()>>>(dsys, s, out_mode, zero ? 1 : 0);
    FPC_LAUNCH_CHECK();
    k_b_reduce<<<dim3(nv, s.n), 1024, 0, stream()>>>(dsys, s, 2, dscal);
    FPC_LAUNCH_CHECK();
    FPC_HIP(hipMemcpyAsync(hscal, dscal, (size_t)s.n * 2 * sizeof(double), hipMemcpyDeviceToHost, stream()));
    FPC_HIP(hipStreamSynchronize(stream()));
    return 0;
  }

  int vcycle(int in_mode, int step, const int *fl) {
    const int nl = (int)down.size();
    for (int l = 0; l < nl; ++l) {
      const Table &dr = dres[(size_t)l], &r = restr[(size_t)l];
      const TileTable &dt = downT[(size_t)l];
      if (dr.nblocks) {
        k_b_down_restrict<<<dr.nblocks, 256, 0, stream()>>>(dlev, dsys, dr.item, dr.first, in_mode, step, fl);
        FPC_LAUNCH_CHECK();
      }
      if (dt.nblocks) {
        k_bt_down<<<dt.nblocks, dim3(kTileX, kTY), 0, stream()>>>(dlev, dsys, dt.ent, in_mode, step, fl);
        FPC_LAUNCH_CHECK();
      }
      if (r.nblocks) {
        k_b_restrict<<<r.nblocks, 256, 0, stream()>>>(dlev, dsys, r.item, r.first, fl);
        FPC_LAUNCH_CHECK();
      }
    }
    k_b_gemv<<<coarse.nblocks, 256, 0, stream()>>>(dlev, dsys, coarse.item, coarse.first, in_mode, step, fl);
    FPC_LAUNCH_CHECK();
    for (int l = nl - 1; l >= 0; --l) {
      const Table &d = down[(size_t)l];
      const TileTable &dt = downT[(size_t)l];
      if (d.nblocks) {
        k_b_up<<<d.nblocks, 256, 0, stream()>>>(dlev, dsys, d.item, d.first, in_mode, step, fl);
        FPC_LAUNCH_CHECK();
      }
      if (dt.nblocks) {
        k_bt_up<<<dt.nblocks, dim3(kTileX, kTY), 0, stream()>>>(dlev, dsys, dt.ent, in_mode, step, fl);
        FPC_LAUNCH_CHECK();
      }
    }
    return 0;
  }

  // The V-cycle of Arnoldi step `step` as a captured graph: captured at the first step
  // with this index and replayed afterwards (the same kernels and arguments; systems
  // already converged skip on their solve state, read when the kernels run).
  int vcycle_graph(int step) {
    hipStream_t st = stream();
    for (const auto &g : vgraphs)
      if (g.first == step) {
        FPC_HIP(hipGraphLaunch(g.second, st));
        return 0;
      }
    FPC_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    const int rc = vcycle(0, step, nullptr);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      if (rc) return rc;
      FPC_HIP(e);
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    FPC_HIP(ei);
    vgraphs.emplace_back(step, exec);
    FPC_HIP(hipGraphLaunch(exec, st));
    return 0;
  }

  // One Arnoldi step of every system still in its cycle; step = host-side basis index.
  int run_step(int step) {
    if (pc == 3) {
      const bool graph = tile_tuning().mg_graph && !prof_enabled();
      int rc = graph ? vcycle_graph(step) : vcycle(0, step, nullptr);
      if (rc) return rc;
    }
    if (pts.nblocks) {
      k_b_matvec<<<pts.nblocks, 256, 0, stream()>>>(dsys, pts.item, pts.first, pc == 3 ? 1 : 0);
      FPC_LAUNCH_CHECK();
    }
    if (ptsT.nblocks) {
      if (pc == 1)
        k_bt_matvec<true><<<ptsT.nblocks, dim3(kTileX, kTY), 0, stream()>>>(dsys, ptsT.ent, 0);
      else
        k_bt_matvec<false><<<ptsT.nblocks, dim3(kTileX, kTY), 0, stream()>>>(dsys, ptsT.ent, pc == 3 ? 1 : 0);
      FPC_LAUNCH_CHECK();
    }
    k_b_multidot<<<dim3(maxnred, n), 256, 0, stream()>>>(dsys);
    FPC_LAUNCH_CHECK();
    k_b_reduce_h<<<dim3(std::min(step + 1, KM + 1), n), 256, 0, stream()>>>(dsys);
    FPC_LAUNCH_CHECK();
    k_b_orth_update<<<dim3(maxnred, n), 256, 0, stream()>>>(dsys);
    FPC_LAUNCH_CHECK();
    k_b_tail<<<n, 256, 0, stream()>>>(dsys);
    FPC_LAUNCH_CHECK();
    return 0;
  }

  int solve(int *its_out, double *rel_out, int *err_out);

  // (Re)builds the launch tables of the matvec and V-cycle levels for tile configuration cfg.
  int build_tables(const TileConfig &cfg) {
    free_graphs();
    free_table(pts);
    free_tiles(ptsT);
    for (auto &t : down) free_table(t);
    for (auto &t : dres) free_table(t);
    for (auto &t : restr) free_table(t);
    for (auto &t : downT) free_tiles(t);
    tiles = cfg;
    std::vector<int> one, til, tnx, tny, tnz;
    std::vector<long long> onen;
    for (int q = 0; q < n; ++q) {
      const BSys &e = hsys[(size_t)q];
      if (tiled_box(e.nx, e.ny, e.N, cfg.tile_min)) {
        til.push_back(q);
        tnx.push_back(e.nx);
        tny.push_back(e.ny);
        tnz.push_back(e.nz);
      } else {
        one.push_back(q);
        onen.push_back(e.N);
      }
    }
    int rc = make_table(one, onen, 256, n, pts);
    if (!rc) rc = make_tile_table(til, tnx, tny, tnz, cfg.kz, ptsT);
    if (rc || pc != 3) return rc;
    const int maxl = *std::max_element(nlev.begin(), nlev.end());
    down.assign((size_t)std::max(0, maxl - 1), Table{});
    dres.assign(down.size(), Table{});
    restr.assign(down.size(), Table{});
    downT.assign(down.size(), TileTable{});
    for (int l = 0; l + 1 < maxl && !rc; ++l) {
      // fine level l of every system with a level below it: one-point (residual and
      // restriction in one launch over the coarse points; the up-sweep over the
      // fine ones) or tiled (the z-coarsened ones restrict inside their down kernel)
      std::vector<int> onel, tl, rit, lnx, lny, lnz;
      std::vector<long long> fine, coarse_one, crs;
      for (int q = 0; q < n; ++q) {
        if (nlev[(size_t)q] <= l + 1) continue;
        const int it = base[(size_t)q] + l;
        const BLev &e = hlev[(size_t)it];
        const bool tiled = tiled_box(e.nx, e.ny, e.N, cfg.tile_min);
        if (tiled) {
          tl.push_back(it);
          lnx.push_back(e.nx);
          lny.push_back(e.ny);
          lnz.push_back(e.nz);
        } else {
          onel.push_back(it);
          fine.push_back(e.N);
          coarse_one.push_back(e.Nc);
        }
        if (tiled && e.cdir != 2) {
          rit.push_back(it);
          crs.push_back(e.Nc);
        }
      }
      const int nids = (int)hlev.size();
      rc = make_table(onel, fine, 256, nids, down[(size_t)l]);
      if (!rc) rc = make_table(onel, coarse_one, 256, nids, dres[(size_t)l]);
      if (!rc) rc = make_tile_table(tl, lnx, lny, lnz, cfg.kz, downT[(size_t)l]);
      if (!rc) rc = make_table(rit, crs, 256, nids, restr[(size_t)l]);
    }
    return rc;
  }

  // Autotuner: times one V-cycle of every system (input W = 0) under each tile
  // configuration and keeps the fastest (all give the same bits).  The V-cycle's
  // buffers hold scratch results afterwards; every solve overwrites them.
  int autotune() {
    if (pc != 3) return 0;
    for (const BSys &e : hsys) FPC_HIP(hipMemsetAsync(e.W, 0, (size_t)e.N * sizeof(double), stream()));
    Sub all;
    all.n = n;
    for (int q = 0; q < n; ++q) all.idx[q] = (unsigned short)q;
    k_b_flags<<<1, 1024, 0, stream()>>>(all, n, flags);
    FPC_LAUNCH_CHECK();
    hipEvent_t e0, e1;
    FPC_HIP(hipEventCreate(&e0));
    FPC_HIP(hipEventCreate(&e1));
    float best = 0.0f;
    TileConfig pick = kTileConfigs[0];
    int rc = 0;
    for (const TileConfig &cfg : kTileConfigs) {
      if ((rc = build_tables(cfg)) || (rc = vcycle(1, 0, flags))) break;  // warm-up
      if (hipEventRecord(e0, stream()) != hipSuccess) { rc = FPC_ERROR_GENERIC; break; }
      for (int r = 0; r < 3 && !rc; ++r) rc = vcycle(1, 0, flags);
      if (rc || hipEventRecord(e1, stream()) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
        rc = rc ? rc : FPC_ERROR_GENERIC;
        break;
      }
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (best == 0.0f || ms < best) {
        best = ms;
        pick = cfg;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    return build_tables(pick);
  }
};

int batch_group_create(const std::vector<fpc_StructSolver> &solvers, const std::vector<const double *> &b,
                       const std::vector<double *> &x, BatchGroup **out) {
  const int n = (int)solvers.size();
  FPC_CHECK_ARG(n >= 1 && n <= kBMax, "a batch group holds 1..768 systems");
  FPC_CHECK_ARG(b.size() == solvers.size() && x.size() == solvers.size(), "size mismatch");
  BatchGroup *G = new BatchGroup();
  G->n = n;
  G->view.resize((size_t)n);
  std::vector<BSys> hs((size_t)n);
  int rc = 0;
  for (int q = 0; q < n && !rc; ++q) {
    GmresBatchView &v = G->view[(size_t)q];
    rc = gmres_batch_view(solvers[(size_t)q], &v);
    if (rc) break;
    if (q == 0) {
      G->pc = v.pc;
      G->flex = v.flexible;
    }
    if (v.pc != G->pc || v.flexible != G->flex) {
      set_error(FPC_ERROR_ARG, "batched solve: all systems of a group need the same preconditioner and GMRES form");
      rc = FPC_ERROR_ARG;
      break;
    }
    Dims d = dims_of(v.A->grid);
    BSys &s = hs[(size_t)q];
    const Planes<double> P = planes64(v.A);
    s.c = P.c;
    s.pitch = P.pitch;
    s.zm = P.zm;
    s.fc = v.A->coef;
    s.fpitch = v.A->pitch;
    s.nx = d.nx;
    s.ny = d.ny;
    s.nz = d.nz;
    s.nred = v.nred;
    s.N = d.N;
    s.vp = v.vp;
    s.dinv = v.dinv ? jacobi_dinv<double>(v.A) : nullptr;
    s.dinv_full = v.dinv;
    s.V = v.V;
    s.Z = v.Z;
    s.flex = v.flexible;
    s.W = v.W;
    s.part = v.part;
    s.S = v.S;
    s.b = b[(size_t)q];
    s.x = x[(size_t)q];
    G->maxnred = std::max(G->maxnred, v.nred);
    G->maxN = std::max(G->maxN, d.N);
  }
  if (!rc) rc = dupload(&G->dsys, hs);
  G->hsys = hs;
  if (!rc && G->pc == 3) {
    // level descriptors: system q's levels are contiguous, coarsest last
    std::vector<std::vector<MgLevelView>> lv((size_t)n);
    std::vector<int> base((size_t)n);
    std::vector<const double *> inv((size_t)n);
    std::vector<int> nco((size_t)n);
    int maxl = 0;
    for (int q = 0; q < n && !rc; ++q) {
      rc = pfmg_batch_view(G->view[(size_t)q].pfmg, lv[(size_t)q], &inv[(size_t)q], &nco[(size_t)q]);
      maxl = std::max(maxl, (int)lv[(size_t)q].size());
    }
    for (int q = 0; q < n && !rc; ++q) {
      base[(size_t)q] = (int)G->hlev.size();
      const auto &L = lv[(size_t)q];
      const int nl = (int)L.size();
      for (int l = 0; l < nl; ++l) {
        BLev e{};
        e.c = L[(size_t)l].coef;
        e.dinv = L[(size_t)l].dinv;
        e.wm = L[(size_t)l].wm;
        e.wp = L[(size_t)l].wp;
        e.pitch = L[(size_t)l].pitch;
        e.zm = L[(size_t)l].zm;
        e.fc = L[(size_t)l].fcoef;
        e.fdinv = L[(size_t)l].fdinv;
        e.fwm = L[(size_t)l].fwm;
        e.fwp = L[(size_t)l].fwp;
        e.fpitch = L[(size_t)l].fpitch;
        e.b = L[(size_t)l].b;
        e.r = L[(size_t)l].r;
        e.y = L[(size_t)l].y;
        e.nx = L[(size_t)l].nx;
        e.ny = L[(size_t)l].ny;
        e.nz = L[(size_t)l].nz;
        e.N = L[(size_t)l].N;
        e.cdir = L[(size_t)l].cdir;
        e.w = L[(size_t)l].w;
        e.sys = q;
        e.lvl = l;
        if (l + 1 < nl) {
          const MgLevelView &C = L[(size_t)l + 1];
          e.bc = C.b;
          e.yc = C.y;
          e.cnx = C.nx;
          e.cny = C.ny;
          e.cnz = C.nz;
          e.Nc = C.N;
          e.ncd = (e.cdir == 0) ? C.nx : (e.cdir == 1 ? C.ny : C.nz);
        } else {
          e.inv = inv[(size_t)q];
          e.ncoarse = nco[(size_t)q];
        }
        G->hlev.push_back(e);
      }
    }
    if (!rc) rc = dupload(&G->dlev, G->hlev);
    G->base = base;
    for (int q = 0; q < n; ++q) G->nlev.push_back((int)lv[(size_t)q].size());
    // coarsest levels: ceil(ncoarse / 4) blocks each
    if (!rc) {
      std::vector<int> items;
      std::vector<long long> rows;
      for (int q = 0; q < n; ++q) {
        items.push_back(base[(size_t)q] + (int)lv[(size_t)q].size() - 1);
        rows.push_back(nco[(size_t)q]);
      }
      rc = make_table(items, rows, 4, (int)G->hlev.size(), G->coarse);
    }
  }
  if (!rc) {
    if (hipMalloc(&G->flags, (size_t)n * sizeof(int)) != hipSuccess || hipMalloc(&G->dst, (size_t)n * sizeof(Status)) ||
        hipMalloc(&G->dscal, (size_t)n * 2 * sizeof(double)) || hipMalloc(&G->dargs, (size_t)n * 5 * sizeof(double)) ||
        hipHostMalloc(&G->hst, (size_t)n * sizeof(Status)) || hipHostMalloc(&G->hscal, (size_t)n * 2 * sizeof(double)) ||
        hipHostMalloc(&G->hargs, (size_t)n * 5 * sizeof(double))) {
      set_error(FPC_ERROR_MEMORY, "batched solve: allocation failed");
      rc = FPC_ERROR_MEMORY;
    }
  }
  if (!rc) rc = G->autotune();
  if (!rc && G->pc != 3) rc = G->build_tables(G->tiles);
  if (!rc) {
    // every state starts finished, so only systems in a cycle run Arnoldi steps
    k_b_set_done<<<(n + 255) / 256, 256, 0, stream()>>>(G->dsys, n, 3);
    rc = (hipGetLastError() == hipSuccess && hipStreamSynchronize(stream()) == hipSuccess) ? 0 : FPC_ERROR_GENERIC;
  }
  if (rc) {
    delete G;
    return rc;
  }
  *out = G;
  return 0;
}

// The control flow of fpc_StructGMRESSolve (krylov.hip) for every system of the
// group at once.
int BatchGroup::solve(int *its_out, double *rel_out, int *err_out) {
  std::vector<double> r_norm((size_t)n), b_norm((size_t)n), eps((size_t)n), real_old((size_t)n);
  std::vector<int> iter((size_t)n, 0), active((size_t)n, 1), first((size_t)n, 1);
  std::vector<int> all((size_t)n);
  for (int q = 0; q < n; ++q) all[(size_t)q] = q;
  // every system hinted to start from x = 0: the initial residuals are copies of b
  bool zero = true;
  for (int q = 0; q < n; ++q) zero = zero && view[(size_t)q].x0_zero && *view[(size_t)q].x0_zero;
  for (int q = 0; q < n; ++q)
    if (view[(size_t)q].x0_zero) *view[(size_t)q].x0_zero = false;
  int rc = true_residual(all, 0, true, zero);
  if (rc) return rc;
  for (int q = 0; q < n; ++q) {
    r_norm[(size_t)q] = std::sqrt(hscal[2 * q]);
    b_norm[(size_t)q] = std::sqrt(hscal[2 * q + 1]);
    *view[(size_t)q].its = 0;
    if (!std::isfinite(b_norm[(size_t)q]) || !std::isfinite(r_norm[(size_t)q])) {
      *view[(size_t)q].relres = NAN;
      set_error(FPC_ERROR_GENERIC, "GMRES: non-finite right-hand side or initial residual");
      return FPC_ERROR_GENERIC;
    }
    const double den = (b_norm[(size_t)q] > 0.0) ? b_norm[(size_t)q] : r_norm[(size_t)q];
    eps[(size_t)q] = std::max(view[(size_t)q].atol, view[(size_t)q].tol * den);
    real_old[(size_t)q] = b_norm[(size_t)q];
  }
  for (;;) {
    std::vector<int> pre, cyc;
    for (int q = 0; q < n; ++q) {
      if (!active[(size_t)q]) continue;
      if (!(iter[(size_t)q] < view[(size_t)q].maxit) || r_norm[(size_t)q] == 0.0) {
        active[(size_t)q] = 0;
        continue;
      }
      (r_norm[(size_t)q] <= eps[(size_t)q] ? pre : cyc).push_back(q);
    }
    if (!pre.empty()) {
      if ((rc = true_residual(pre, 1, false))) return rc;
      std::vector<int> cp;
      for (size_t e = 0; e < pre.size(); ++e) {
        const int q = pre[e];
        r_norm[(size_t)q] = std::sqrt(hscal[2 * e]);
        if (r_norm[(size_t)q] <= eps[(size_t)q]) {
          active[(size_t)q] = 0;
        } else {
          cp.push_back(q);
          cyc.push_back(q);
        }
      }
      if (!cp.empty()) {
        const Sub s = sub_of(cp);
        k_b_copy_w_v0<<<dim3(grid_blocks(maxN), s.n), 256, 0, stream()>>>(dsys, s);
        FPC_LAUNCH_CHECK();
      }
    }
    if (cyc.empty()) break;
    std::sort(cyc.begin(), cyc.end());
    const Sub sc = sub_of(cyc);
    int kdmax = 0, batch = 1, room = 1;
    for (int e = 0; e < sc.n; ++e) {
      const int q = cyc[(size_t)e];
      const GmresBatchView &v = view[(size_t)q];
      double *a = hargs + 5 * e;
      a[0] = r_norm[(size_t)q];
      a[1] = eps[(size_t)q];
      a[2] = iter[(size_t)q];
      a[3] = v.maxit;
      a[4] = v.kd;
      kdmax = std::max(kdmax, v.kd);
      batch = std::max(batch, std::min(*v.pred, v.kd));
      room = std::max(room, v.maxit - iter[(size_t)q]);
    }
    FPC_HIP(hipMemcpyAsync(dargs, hargs, (size_t)sc.n * 5 * sizeof(double), hipMemcpyHostToDevice, stream()));
    k_b_cycle_start<<<(sc.n + 255) / 256, 256, 0, stream()>>>(dsys, sc, dargs);
    FPC_LAUNCH_CHECK();
    // Arnoldi steps in batches until every system of the cycle has stopped
    int issued = 0;
    for (;;) {
      int nb = std::min(batch, kdmax - issued);
      nb = std::min(nb, room - issued);
      if (nb <= 0) nb = 1;
      for (int q = 0; q < nb; ++q)
        if ((rc = run_step(issued + q))) return rc;
      issued += nb;
      k_b_status<<<(n + 255) / 256, 256, 0, stream()>>>(dsys, n, dst);
      FPC_LAUNCH_CHECK();
      FPC_HIP(hipMemcpyAsync(hst, dst, (size_t)n * sizeof(Status), hipMemcpyDeviceToHost, stream()));
      FPC_HIP(hipStreamSynchronize(stream()));
      bool all_done = true;
      for (int q : cyc) all_done = all_done && hst[q].done != 0;
      if (all_done || issued >= kdmax) break;
      batch = (pc == 3) ? 1 : 2;
    }
    for (int q : cyc) {
      const Status &st = hst[q];
      iter[(size_t)q] = st.iter;
      r_norm[(size_t)q] = st.r_norm;
      if (first[(size_t)q]) *view[(size_t)q].pred = std::max(1, st.i);
      first[(size_t)q] = 0;
    }
    // x update
    k_b_backsolve<<<(sc.n + 255) / 256, 256, 0, stream()>>>(dsys, sc);
    FPC_LAUNCH_CHECK();
    const unsigned gb = grid_blocks(maxN);
    if (pc == 3 && flex) {
      // flexible: the preconditioned vectors are stored; no V-cycle of the combination
      k_b_combine<<<dim3(gb, sc.n), 256, 0, stream()>>>(dsys, sc, 2);
      FPC_LAUNCH_CHECK();
    } else if (pc == 3) {
      k_b_combine<<<dim3(gb, sc.n), 256, 0, stream()>>>(dsys, sc, 1);
      FPC_LAUNCH_CHECK();
      k_b_flags<<<1, 1024, 0, stream()>>>(sc, n, flags);
      FPC_LAUNCH_CHECK();
      if ((rc = vcycle(1, 0, flags))) return rc;
      k_b_axpy_z<<<dim3(gb, sc.n), 256, 0, stream()>>>(dsys, sc);
      FPC_LAUNCH_CHECK();
    } else {
      k_b_combine<<<dim3(gb, sc.n), 256, 0, stream()>>>(dsys, sc, 0);
      FPC_LAUNCH_CHECK();
    }
    std::vector<int> tr, rs;
    for (int q : cyc) (r_norm[(size_t)q] <= eps[(size_t)q] ? tr : rs).push_back(q);
    if (!tr.empty()) {
      if ((rc = true_residual(tr, 0, false))) return rc;
      for (size_t e = 0; e < tr.size(); ++e) {
        const int q = tr[e];
        const double real_new = std::sqrt(hscal[2 * e]);
        r_norm[(size_t)q] = real_new;
        if (real_new <= eps[(size_t)q] || real_new >= real_old[(size_t)q]) active[(size_t)q] = 0;
        else real_old[(size_t)q] = real_new;
      }
    }
    if (!rs.empty()) {
      const Sub sr = sub_of(rs);
      k_b_restart_coefs<<<(sr.n + 255) / 256, 256, 0, stream()>>>(dsys, sr);
      FPC_LAUNCH_CHECK();
      k_b_restart_vec<<<dim3(gb, sr.n), 256, 0, stream()>>>(dsys, sr);
      FPC_LAUNCH_CHECK();
    }
  }
  FPC_HIP(hipStreamSynchronize(stream()));
  for (int q = 0; q < n; ++q) {
    const GmresBatchView &v = view[(size_t)q];
    *v.its = iter[(size_t)q];
    *v.relres = (b_norm[(size_t)q] > 0.0) ? r_norm[(size_t)q] / b_norm[(size_t)q] : r_norm[(size_t)q];
    if (its_out) its_out[q] = iter[(size_t)q];
    if (rel_out) rel_out[q] = *v.relres;
    if (err_out)
      err_out[q] = (iter[(size_t)q] >= v.maxit && r_norm[(size_t)q] > eps[(size_t)q] && eps[(size_t)q] > 0)
                       ? FPC_ERROR_CONV
                       : 0;
  }
  return 0;
}

int batch_group_solve(BatchGroup *g, int *its, double *rel, int *err) { return g->solve(its, rel, err); }

void batch_group_destroy(BatchGroup *g) { delete g; }

void batch_group_tiles(const BatchGroup *g, int *kz, long long *tile_min) {
  *kz = g->tiles.kz;
  *tile_min = g->tiles.tile_min;
}

// ------------------------------------------------------------ non-finite check
class NonfiniteCheck {
 public:
  int n = 0;
  const double **arr = nullptr;
  long long *len = nullptr;
  int *dflags = nullptr, *hflags = nullptr;
  long long maxlen = 0;
  ~NonfiniteCheck() {
    if (arr) (void)hipFree(arr);
    if (len) (void)hipFree(len);
    if (dflags) (void)hipFree(dflags);
    if (hflags) (void)hipHostFree(hflags);
  }
};

int nonfinite_create(const std::vector<const double *> &arrays, const std::vector<long long> &len,
                     NonfiniteCheck **out) {
  FPC_CHECK_ARG(!arrays.empty() && arrays.size() == len.size() && arrays.size() <= 65535, "bad arrays");
  NonfiniteCheck *c = new NonfiniteCheck();
  c->n = (int)arrays.size();
  for (long long l : len) c->maxlen = std::max(c->maxlen, l);
  std::vector<const double *> a(arrays);
  int rc = dupload(&c->arr, a);
  if (!rc) rc = dupload(&c->len, len);
  if (!rc && (hipMalloc(&c->dflags, (size_t)c->n * sizeof(int)) != hipSuccess ||
              hipHostMalloc(&c->hflags, (size_t)c->n * sizeof(int)) != hipSuccess))
    rc = FPC_ERROR_MEMORY;
  if (rc) {
    delete c;
    return rc;
  }
  *out = c;
  return 0;
}

int nonfinite_enqueue(NonfiniteCheck *c) {
  k_b_zero_int<<<(c->n + 255) / 256, 256, 0, stream()>>>(c->dflags, c->n);
  FPC_LAUNCH_CHECK();
  k_b_nonfinite<<<dim3(grid_blocks(c->maxlen), c->n), 256, 0, stream()>>>(c->arr, c->len, c->dflags);
  FPC_LAUNCH_CHECK();
  FPC_HIP(hipMemcpyAsync(c->hflags, c->dflags, (size_t)c->n * sizeof(int), hipMemcpyDeviceToHost, stream()));
  return 0;
}

const int *nonfinite_flags(NonfiniteCheck *c) { return c->hflags; }

void nonfinite_destroy(NonfiniteCheck *c) { delete c; }

}  // namespace fpc

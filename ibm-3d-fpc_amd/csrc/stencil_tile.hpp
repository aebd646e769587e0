// stencil_tile.hpp -- z-marching tiles for the 7-point box stencils (libfpc).
//
// A workgroup of TX x TY threads owns an x-y tile and walks a chunk of z-planes.
// Each thread keeps the stencil input of its own column at k-1, k, k+1 in
// registers, so the z-neighbours cost no extra memory traffic; the x/y
// neighbours come from the current plane staged in LDS (double-buffered, one
// barrier per plane).  The ring of 2*TX + 2*TY halo points around the tile is
// evaluated by the first threads of the workgroup.  The stencil input is any
// point function ("val"): a plain vector load, v * (1/diag), a Jacobi sweep
// from zero, or an interpolated coarse-grid correction, so the same walk serves
// SpMV, the MG smoothers and the Arnoldi matvec.
//
// Every global row is read once (x-contiguous, 64 lanes = one 512-byte FP64
// row segment); the only redundant work is the halo ring (2/TY + 2/TX of the
// point evaluations, served from L2) and two extra planes per z-chunk.
#ifndef FPC_STENCIL_TILE_HPP
#define FPC_STENCIL_TILE_HPP

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include "fpc_internal.hpp"

namespace fpc {

constexpr int kTileX = 64;
// Longest z-chunk of a marching tile: its fetches then span at most two 64-plane flag
// words per column (march_tile).
constexpr int kMaxChunk = 62;

// Halo ring point q of a TX x TY tile -> offset (ox, oy) in [-1, TX] x [-1, TY].
template <int TX, int TY>
__device__ inline bool tile_ring(int q, int &ox, int &oy) {
  if (q < TX) { ox = q; oy = -1; return true; }
  q -= TX;
  if (q < TX) { ox = q; oy = TY; return true; }
  q -= TX;
  if (q < TY) { ox = -1; oy = q; return true; }
  q -= TY;
  if (q < TY) { ox = TX; oy = q; return true; }
  return false;
}

// Neighbour values of one point, in the stencil's slot order after the centre;
// bit s-1 of in_box is set when neighbour s (-x, +x, -y, +y, -z, +z) lies in
// the box.
template <typename T>
struct Nbr {
  T c, xm, xp, ym, yp, zm, zp;
  unsigned in_box;
};

__device__ inline unsigned nbr_in_box(int i, int j, int k, int nx, int ny, int nz) {
  return (unsigned)(i > 0) | ((unsigned)(i < nx - 1) << 1) | ((unsigned)(j > 0) << 2) | ((unsigned)(j < ny - 1) << 3) |
         ((unsigned)(k > 0) << 4) | ((unsigned)(k < nz - 1) << 5);
}

// The seven coefficients of a row; apply() sums in slot order C, -x, +x, -y,
// +y, -z, +z over the neighbours in the box.  Terms of absent neighbours are
// skipped rather than added as a*0, so the sum is bit for bit that of the
// one-point-per-thread kernels (a*0 would turn a -0 partial sum into +0).
template <typename T>
struct Coef7 {
  T a[7];
  __device__ inline void load(const T *c, long long pitch, long long t) {
#pragma unroll
    for (int s = 0; s < 7; ++s) a[s] = c[s * pitch + t];
  }
  __device__ inline T apply(const Nbr<T> &nb) const {
    T s = a[0] * nb.c;
    if (nb.in_box & 1u) s += a[1] * nb.xm;
    if (nb.in_box & 2u) s += a[2] * nb.xp;
    if (nb.in_box & 4u) s += a[3] * nb.ym;
    if (nb.in_box & 8u) s += a[4] * nb.yp;
    if (nb.in_box & 16u) s += a[5] * nb.zm;
    if (nb.in_box & 32u) s += a[6] * nb.zp;
    return s;
  }
};

// Walk planes k0 .. kend-1 of tile (bx, by), owned by this workgroup.  The op
// splits its point function so loads run ahead of use: op.fetch(i, j, k, f) issues
// a point's loads given its hot-copy flag f (an Op::Raw, with Raw::u the row of its
// coefficients) and
// op.finish(raw) gives the value its neighbours see; op.out(a, own, i, j, k, nb)
// writes the point from its coefficients a, its own raw and the neighbour values.
// The own column is fetched two planes ahead and the halo ring one plane ahead
// (three and two for ops with Op::kAhead = 2);
// a barrier waits only for LDS (s_waitcnt lgkmcnt), so those loads stay in flight
// across it.  The coefficients stay in registers while the row does not change
// -- the interior planes of a z-extruded operator's hot copy share one row -- so
// in steady state no coefficient is loaded at all; a point whose row changes
// (wall planes, flagged points, full storage) loads them at the top of the plane.
// Points outside the box read as zero.
template <typename T, int TX, int TY, class Op>
__device__ inline void march_tile(Op &op, int nx, int ny, int nz, int k0, int kend, int bx, int by) {
  using Raw = typename Op::Raw;
  __shared__ T plane[2][TY + 2][TX + 2];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i = bx * TX + tx, j = by * TY + ty;
  const bool in = (i < nx) && (j < ny);
  int hx = 0, hy = 0;
  const bool ring = tile_ring<TX, TY>(ty * TX + tx, hx, hy);
  const int hi = bx * TX + hx, hj = by * TY + hy;
  const bool ring_in = ring && hi >= 0 && hi < nx && hj >= 0 && hj < ny;
  // The own and ring columns' flag words of a hot copy (bit k & 63 of word k >> 6; 0
  // without flags).  A chunk's fetches cover planes k0-1 .. kend, at most 65 planes
  // (chunks are at most kMaxChunk planes), so two words: both loaded here, and a
  // fetch takes its point's flag from them by a select -- no load per plane, whose
  // wait would also wait for every prefetch in flight.
  const int w0 = (k0 > 0 ? k0 - 1 : 0) >> 6, w1 = (w0 + 1) * 64 < nz ? w0 + 1 : w0;
  const long long co = in ? (long long)j * nx + i : 0, cr = ring_in ? (long long)hj * nx + hi : 0;
  const unsigned long long o0 = op.zm.word(co, w0), o1 = op.zm.word(co, w1);
  const unsigned long long r0 = op.zm.word(cr, w0), r1 = op.zm.word(cr, w1);
  pin_all(o0, o1, r0, r1);  // their wait here, once, not at every use in the loop
  auto fown = [&](int kk) -> int { return (int)((((kk >> 6) == w0 ? o0 : o1) >> (kk & 63)) & 1ull); };
  auto fring = [&](int kk) -> int { return (int)((((kk >> 6) == w0 ? r0 : r1) >> (kk & 63)) & 1ull); };
  T zm = (in && k0 > 0) ? op.finish(op.fetch(i, j, k0 - 1, fown(k0 - 1))) : T(0);
  // Prefetch distance: the own column D+1 planes and the ring D planes ahead.  Op::kAhead
  // is 2 for ops whose raws are small enough to keep one more plane of each in flight
  // at the same occupancy (the down sweeps), else 1.
  constexpr int D = Op::kAhead;
  static_assert(D == 1 || D == 2, "prefetch distance 1 or 2");
  Raw own[D + 2]{}, ringq[D + 1]{};
#pragma unroll
  for (int q = 0; q <= D; ++q)
    if (in && k0 + q < nz && (q < 2 || k0 + q - 1 < kend)) own[q] = op.fetch(i, j, k0 + q, fown(k0 + q));
#pragma unroll
  for (int q = 0; q < D; ++q)
    if (ring_in && k0 + q < kend) ringq[q] = op.fetch(hi, hj, k0 + q, fring(k0 + q));
  Raw &own0 = own[0], &own1 = own[1];
  T zc = (in && k0 < nz) ? op.finish(own0) : T(0);
  Coef7<T> a;
  long long arow = -1;
  for (int k = k0; k < kend; ++k) {
    if (in && own0.u != arow) {
      a.load(op.c, op.pitch, own0.u);
      arow = own0.u;
    }
    own[D + 1] = Raw{};
    ringq[D] = Raw{};
    if (in && k + D + 1 < nz && k + D < kend) own[D + 1] = op.fetch(i, j, k + D + 1, fown(k + D + 1));
    if (ring_in && k + D < kend) ringq[D] = op.fetch(hi, hj, k + D, fring(k + D));
    Raw &ring0 = ringq[0];
    const T zp = (in && k + 1 < nz) ? op.finish(own1) : T(0);
    T(*P)[TX + 2] = plane[(k - k0) & 1];
    P[ty + 1][tx + 1] = zc;
    if (ring) P[hy + 1][hx + 1] = ring_in ? op.finish(ring0) : T(0);
    __syncthreads();
    if (in) {
      Nbr<T> nb{zc, P[ty + 1][tx], P[ty + 1][tx + 2], P[ty][tx + 1], P[ty + 2][tx + 1], zm, zp,
                nbr_in_box(i, j, k, nx, ny, nz)};
      op.out(a, own0, i, j, k, nb);
    }
    zm = zc;
    zc = zp;
#pragma unroll
    for (int q = 0; q <= D; ++q) own[q] = own[q + 1];
#pragma unroll
    for (int q = 0; q < D; ++q) ringq[q] = ringq[q + 1];
  }
}

// The same walk, software-pipelined (used by the MG down sweeps, whose raws are small):
// the own column is fetched D+1 planes ahead and the halo ring D planes ahead
// (D = Op::kAheadPipelined) into a ring of P = D+2 raws per stream, and the plane loop
// is unrolled by P, so every raw slot is a fixed register set.  march_tile rotates its
// raws from register to register at the end of each plane, and a move of a register
// whose load is in flight waits for it: there every plane waits for all its loads
// (s_waitcnt vmcnt(0) before the barrier).  Here each use waits only for its own load
// (vmcnt(n), n = loads issued after it), so the loads of the next planes stay in flight
// across the barrier.  Every fetch is unconditional: a lane outside the box or the
// ring, or a plane past the walk, loads a valid point already read (the nearest box
// point, the last plane the walk needs), which keeps the count of loads in flight the
// same on every lane and path.  The values computed are those of march_tile, bit for
// bit.  The unrolled slots cost registers (126 VGPRs for the z-restricting down sweep
// at 64x4 tiles and D = 1), so the up sweeps, whose raws are larger, keep march_tile.
template <typename T, int TX, int TY, class Op>
__device__ inline void march_tile_pipelined(Op &op, int nx, int ny, int nz, int k0, int kend, int bx, int by) {
  using Raw = typename Op::Raw;
  constexpr int D = Op::kAheadPipelined;
  static_assert(D == 1 || D == 2, "prefetch distance 1 or 2");
  constexpr int P = D + 2;
  __shared__ T plane[2][TY + 2][TX + 2];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i = bx * TX + tx, j = by * TY + ty;
  const bool in = (i < nx) && (j < ny);
  const int ic = i < nx ? i : nx - 1, jc = j < ny ? j : ny - 1;
  int hx = 0, hy = 0;
  const bool ring = tile_ring<TX, TY>(ty * TX + tx, hx, hy);
  const int hi = bx * TX + hx, hj = by * TY + hy;
  const bool ring_in = ring && hi >= 0 && hi < nx && hj >= 0 && hj < ny;
  const int hic = hi < 0 ? 0 : (hi < nx ? hi : nx - 1), hjc = hj < 0 ? 0 : (hj < ny ? hj : ny - 1);
  // The own and ring columns' flag words of a hot copy (bit k & 63 of word k >> 6; 0
  // without flags).  A chunk's fetches cover planes k0-1 .. kend, at most 65 planes
  // (chunks are at most kMaxChunk planes), so two words: both loaded here, and a
  // fetch takes its point's flag from them by a select -- no load per plane.
  const int w0 = (k0 > 0 ? k0 - 1 : 0) >> 6, w1 = (w0 + 1) * 64 < nz ? w0 + 1 : w0;
  const long long co = (long long)jc * nx + ic, cr = (long long)hjc * nx + hic;
  const unsigned long long o0 = op.zm.word(co, w0), o1 = op.zm.word(co, w1);
  const unsigned long long r0 = op.zm.word(cr, w0), r1 = op.zm.word(cr, w1);
  pin_all(o0, o1, r0, r1);  // their wait here, once, not at every use in the loop
  auto fown = [&](int kk) -> int { return (int)((((kk >> 6) == w0 ? o0 : o1) >> (kk & 63)) & 1ull); };
  auto fring = [&](int kk) -> int { return (int)((((kk >> 6) == w0 ? r0 : r1) >> (kk & 63)) & 1ull); };
  // planes the walk needs: own column k0-1 .. min(kend, nz-1), ring k0 .. kend-1
  const int ko_last = kend < nz ? kend : nz - 1, kr_last = kend - 1;
  auto fo = [&](int kk) -> Raw {
    kk = kk < ko_last ? kk : ko_last;
    return op.fetch(ic, jc, kk, fown(kk));
  };
  auto fr = [&](int kk) -> Raw {
    kk = kk < kr_last ? kk : kr_last;
    return op.fetch(hic, hjc, kk, fring(kk));
  };
  const Raw below = fo(k0 > 0 ? k0 - 1 : 0);
  Raw own[P], rq[P];
#pragma unroll
  for (int q = 0; q <= D; ++q) own[q] = fo(k0 + q);
#pragma unroll
  for (int q = 0; q < D; ++q) rq[q] = fr(k0 + q);
  T zm = (in && k0 > 0) ? op.finish(below) : T(0);
  T zc = in ? op.finish(own[0]) : T(0);
  Coef7<T> a;
  long long arow = -1;
  // plane k in slot s = (k - k0) mod P
  auto plane_step = [&](const int s, const int k) {
    if (in && own[s].u != arow) {
      a.load(op.c, op.pitch, own[s].u);
      arow = own[s].u;
    }
    own[(s + D + 1) % P] = fo(k + D + 1);
    rq[(s + D) % P] = fr(k + D);
    const T zp = (in && k + 1 < nz) ? op.finish(own[(s + 1) % P]) : T(0);
    T(*Pl)[TX + 2] = plane[(k - k0) & 1];
    Pl[ty + 1][tx + 1] = zc;
    if (ring) Pl[hy + 1][hx + 1] = ring_in ? op.finish(rq[s]) : T(0);
    __syncthreads();
    if (in) {
      Nbr<T> nb{zc, Pl[ty + 1][tx], Pl[ty + 1][tx + 2], Pl[ty][tx + 1], Pl[ty + 2][tx + 1], zm, zp,
                nbr_in_box(i, j, k, nx, ny, nz)};
      op.out(a, own[s], i, j, k, nb);
    }
    zm = zc;
    zc = zp;
  };
  int kb = k0;
  for (; kb + P <= kend; kb += P) {
#pragma unroll
    for (int s = 0; s < P; ++s) plane_step(s, kb + s);
  }
#pragma unroll
  for (int s = 0; s < P - 1; ++s)
    if (kb + s < kend) plane_step(s, kb + s);
}

// Launch geometry of a z-marching kernel: tile rows TY and z-chunk length.
struct TileLaunch {
  dim3 grid, block;
  int kz;
};

// Launch tuning (fpc_SetTuning; environment defaults FPC_TILE_MIN_POINTS and
// FPC_STENCIL_KZ).  0 for kz = automatic.
struct TileTuning {
  long long min_points = -1;
  int kz = -1;
  int z_extrude = 1;  // three-plane copies of z-extruded operators
  int mg_wide_finest = 1;  // FP32 V-cycles read and write the FP64 vectors on the finest level
  int p_flexible = 1;      // the time step's pressure GMRES keeps its preconditioned vectors
  int mg_graph = 0;        // V-cycles replayed as captured graphs
  int mg_block = 64;       // workgroup size of the one-point multigrid kernels
};
inline TileTuning &tile_tuning() {
  static TileTuning t = [] {
    TileTuning x;
    const char *e = std::getenv("FPC_TILE_MIN_POINTS");
    x.min_points = e ? std::max(0LL, std::atoll(e)) : (1LL << 19);
    e = std::getenv("FPC_STENCIL_KZ");
    x.kz = e ? std::max(0, std::atoi(e)) : 0;
    e = std::getenv("FPC_Z_EXTRUDE");
    x.z_extrude = e ? (std::atoi(e) != 0) : 1;
    e = std::getenv("FPC_MG_WIDE_FINEST");
    x.mg_wide_finest = e ? (std::atoi(e) != 0) : 1;
    e = std::getenv("FPC_P_FLEXIBLE");
    x.p_flexible = e ? (std::atoi(e) != 0) : 1;
    e = std::getenv("FPC_MG_GRAPH");
    x.mg_graph = e ? (std::atoi(e) != 0) : 0;
    e = std::getenv("FPC_MG_BLOCK");
    x.mg_block = e ? std::atoi(e) : 64;
    if (x.mg_block != 64 && x.mg_block != 128 && x.mg_block != 256) x.mg_block = 64;
    return x;
  }();
  return t;
}

// Smaller boxes run the one-point-per-thread kernels.
inline bool tile_worthwhile(const Dims &d) {
  return d.nx >= 32 && d.ny >= 8 && d.N >= std::max(tile_tuning().min_points, 1LL << 15);
}

// Algorithmic bytes of one stencil pass over A: per_point bytes of vectors at
// every point (plus, for a hot copy with flagged points, its flag words: 8 bytes
// per 64 planes of a column), and the 7 coefficient planes and a dinv_bytes-wide
// 1/diag plane (0: none) over the rows read.
// hot: the pass reads the hot copy when A has one (tiled launches), else the full planes.
inline double stencil_bytes(const fpc_StructMatrix_s *A, int dinv_bytes, double per_point, bool hot) {
  const bool z = hot && A->zmap.pl && tile_worthwhile(dims_of(A->grid));
  const double rows = z ? stored_rows(A->zmap, A->grid.N, A->zexc) : (double)A->grid.N;
  const int nz = A->grid.n[2];
  const double flags = (z && A->zmap.xb) ? 8.0 * ((nz + 63) / 64) / nz : 0.0;
  return (double)A->grid.N * (per_point + flags) + rows * (8.0 * kSlots + dinv_bytes);
}

// z-chunk length: about 2048 workgroups in flight, chunks of at least 4 planes,
// always even (z-restriction pairs stay inside a chunk).
inline int tile_kz(const Dims &d, int ty) {
  int kz = tile_tuning().kz;
  if (kz <= 0) {
    long long tiles = (long long)((d.nx + kTileX - 1) / kTileX) * ((d.ny + ty - 1) / ty);
    long long chunks = std::max(1LL, 2048 / std::max(1LL, tiles));
    kz = std::max((int)((d.nz + chunks - 1) / chunks), 4);
  }
  kz += kz & 1;
  return kz < kMaxChunk ? kz : kMaxChunk;
}

template <int TY>
inline TileLaunch tile_launch(const Dims &d) {
  TileLaunch L;
  L.kz = tile_kz(d, TY);
  L.block = dim3(kTileX, TY, 1);
  L.grid = dim3((unsigned)((d.nx + kTileX - 1) / kTileX), (unsigned)((d.ny + TY - 1) / TY),
                (unsigned)((d.nz + L.kz - 1) / L.kz));
  return L;
}

}  // namespace fpc

#endif

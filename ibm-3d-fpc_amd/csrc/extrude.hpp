// extrude.hpp -- host planning and layout of the hot copies of z-extruded
// operators (libfpc).  See ZMap (fpc_internal.hpp) for how kernels read them.
#pragma once

#include <cstring>
#include <vector>

namespace fpc {

struct ExtrudePlan {
  int m = 0;            // wall planes kept per side; 0: no hot copy
  long long nexc = 0;   // flagged points (rows read from the full part)
  long long xoff = 0;   // first row of the full part
  std::vector<unsigned char> xf;  // per point: 1 = flagged
};

// arrays: the planes that must repeat, each N = nx ny nz values (x fastest, then
// y, then z).  For m = 1..4 (nz >= 2m + 3), a point on planes m+1 .. nz-1-m is
// flagged when some array's value there differs bit for bit from plane m's at
// the same (i, j).  Cost of m: (2m+1) nx ny compact rows plus 8 rows per flagged
// point (its scattered row costs about a sector per plane).  The cheapest m is
// kept when it reads under half of the N rows.
inline ExtrudePlan plan_extrusion(int nx, int ny, int nz, const std::vector<const double *> &arrays) {
  ExtrudePlan best;
  const long long pl = (long long)nx * ny, N = pl * nz;
  double best_cost = 0.5 * (double)N;
  std::vector<unsigned char> xf;
  for (int m = 1; m <= 4 && nz >= 2 * m + 3; ++m) {
    const double cap = (best_cost - (2.0 * m + 1.0) * (double)pl) / 8.0;  // flags this m can afford
    if (cap < 0.0) break;  // a deeper wall layer only costs more
    xf.assign((size_t)N, 0);
    long long nexc = 0;
    for (int k = m + 1; k <= nz - 1 - m && (double)nexc <= cap; ++k)
      for (long long q = 0; q < pl; ++q) {
        const size_t t = (size_t)k * (size_t)pl + (size_t)q, r = (size_t)m * (size_t)pl + (size_t)q;
        bool diff = false;
        for (const double *a : arrays) diff |= std::memcmp(a + t, a + r, sizeof(double)) != 0;
        if (diff) {
          xf[t] = 1;
          ++nexc;
        }
      }
    const double cost = (2.0 * m + 1.0) * (double)pl + 8.0 * (double)nexc;
    if ((double)nexc <= cap && cost < best_cost) {
      best_cost = cost;
      best.m = m;
      best.nexc = nexc;
      best.xf.swap(xf);
    }
  }
  if (best.m) best.xoff = ((2LL * best.m + 1) * pl + 63) / 64 * 64;
  return best;
}

// The flags as bits per column: bit (k & 63) of word (k >> 6) * nx ny + (j nx + i) is
// set when point (i, j, k) is flagged.  A z-marching tile keeps its columns' words in
// registers, so knowing a point's row needs no load of its own.
inline std::vector<unsigned long long> flag_words(const ExtrudePlan &plan, int nx, int ny, int nz) {
  const long long pl = (long long)nx * ny, nw = (nz + 63) / 64;
  std::vector<unsigned long long> w((size_t)(nw * pl), 0ull);
  for (int k = 0; k < nz; ++k)
    for (long long c = 0; c < pl; ++c)
      if (plan.xf[(size_t)(k * pl + c)]) w[(size_t)((k >> 6) * pl + c)] |= 1ull << (k & 63);
  return w;
}

// Source plane of compact plane kc (0 .. 2m).
inline long long hot_source_plane(int kc, int m, int nz) {
  return (kc < m) ? kc : (kc == m ? m : (long long)nz - 2 * m - 1 + kc);
}

// Hot layout of nslots planes (slot s at src + s * spitch): per slot, the 2m+1
// compact planes, then all N rows from xoff on; slots dpitch apart.
inline std::vector<double> hot_planes(const double *src, long long spitch, int nslots, int nx, int ny, int nz,
                                      const ExtrudePlan &p, long long dpitch) {
  const long long pl = (long long)nx * ny, N = pl * nz;
  std::vector<double> out((size_t)nslots * (size_t)dpitch, 0.0);
  for (int s = 0; s < nslots; ++s) {
    const double *a = src + (size_t)s * (size_t)spitch;
    double *o = out.data() + (size_t)s * (size_t)dpitch;
    for (int kc = 0; kc <= 2 * p.m; ++kc)
      std::memcpy(o + (size_t)kc * (size_t)pl, a + (size_t)hot_source_plane(kc, p.m, nz) * (size_t)pl,
                  (size_t)pl * sizeof(double));
    std::memcpy(o + (size_t)p.xoff, a, (size_t)N * sizeof(double));
  }
  return out;
}

}  // namespace fpc

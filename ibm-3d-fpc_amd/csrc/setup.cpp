// setup.cpp -- one-time host setup of the channel operators (libfpc).
//
// Restates the reference's CPU pipeline (immersed_geometry.hpp:1-160,
// main.cu:33-121, immersed_operators.hpp:1-287, 397-447, 737-748,
// solver_helpers.hpp:51-84, 227-329) directly on the structured grid: every
// embedded 7-point row is computed in place, with no sparse maps, in O(N).
//
// Bit-exactness.  The reference accumulates two things through
// std::unordered_map<int,double> rows, whose iteration order then fixes the
// rounding:
//   * the l1 row scale 1/sum|a| of the u, v, w matrices (immersed_operators.hpp:742);
//   * the diagonal of Ap = D*G, summed over D's row in map order (common.hpp:175-189).
// map_order() reproduces that order: libstdc++'s table has 13 buckets once the
// first key is inserted (room for 13 keys, rows here have at most 7), a key
// joins the front of its bucket's run, and a bucket's first key goes to the
// front of the whole list; std::hash<int> is the identity.  The add_entry
// (|v| < 1e-30) and compress_zeros (|v| <= 1e-14) filters are applied at the
// same points as in the reference.
//
// Compiled with -ffp-contract=off: host arithmetic rounds exactly as written.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "fpc_setup.hpp"

namespace fpc {
namespace host {

namespace {

inline size_t uix(int i, int j, int k, int nx, int ny) { return ((size_t)k * ny + j) * (size_t)(nx + 1) + (size_t)i; }
inline size_t vix(int i, int j, int k, int nx, int ny) { return ((size_t)k * (ny + 1) + j) * (size_t)nx + (size_t)i; }
inline size_t wix(int i, int j, int k, int nx, int ny) { return ((size_t)k * ny + j) * (size_t)nx + (size_t)i; }
inline size_t pix(int i, int j, int k, int nx, int ny) { return ((size_t)k * ny + j) * (size_t)nx + (size_t)i; }

// Distance along the segment (xP,yP)->(xS,yS) to the circle, or +inf (common.hpp:262-280).
double ray_to_circle(double xP, double yP, double xS, double yS, const fpc_params &p) {
  const double inf = std::numeric_limits<double>::infinity();
  double ex = xS - xP, ey = yS - yP;
  double a = ex * ex + ey * ey;
  if (a <= 0.0) return inf;
  double ox = xP - p.xcirc, oy = yP - p.ycirc;
  double b = 2.0 * (ex * ox + ey * oy);
  double c = ox * ox + oy * oy - p.Rcirc * p.Rcirc;
  double disc = b * b - 4.0 * a * c;
  if (disc < 0.0) return inf;
  double r = std::sqrt(std::max(0.0, disc));
  double t1 = (-b - r) / (2.0 * a), t2 = (-b + r) / (2.0 * a);
  double t;
  if (t1 >= 0.0 && t1 <= 1.0) t = t1;
  else if (t2 >= 0.0 && t2 <= 1.0) t = t2;
  else return inf;
  return std::sqrt(a) * t;
}

// Iteration order of keys[0..n-1] inserted into an empty std::unordered_map<int,..>.
void map_order(const int *keys, int n, int *order) {
  int len = 0;
  for (int q = 0; q < n; ++q) {
    int b = (int)((unsigned)keys[q] % 13u);
    int pos = 0;
    for (int t = 0; t < len; ++t)
      if ((int)((unsigned)keys[order[t]] % 13u) == b) { pos = t; break; }
    for (int t = len; t > pos; --t) order[t] = order[t - 1];
    order[pos] = q;
    ++len;
  }
}

// One active row of a u/v/w diffusion matrix before scaling: entries in the
// reference's insertion order (W, E, S, N, B, T, P) with their stencil slots.
struct Row {
  int n = 0;
  int key[7];
  int slot[7];
  double val[7];
  void add(int k, int s, double v) {
    if (std::abs(v) < 1.0e-30) return;  // add_entry
    key[n] = k;
    slot[n] = s;
    val[n] = v;
    ++n;
  }
};

// l1 row scaling of one row and its AoS embedding (immersed_operators.hpp:737-748).
double scale_and_embed(const Row &r, double *a7) {
  int ord[7];
  map_order(r.key, r.n, ord);
  double s = 0.0;
  for (int t = 0; t < r.n; ++t) {
    double v = r.val[ord[t]];
    if (std::abs(v) <= 1.0e-14) continue;  // compress_zeros of the unscaled matrix
    s += std::abs(v);
  }
  if (s < 1.0e-14) s = 1.0;
  double sc = 1.0 / s;
  for (int t = 0; t < r.n; ++t) {
    double v = r.val[t];
    if (std::abs(v) <= 1.0e-14) continue;
    double w = sc * v;
    if (std::abs(w) < 1.0e-30 || std::abs(w) <= 1.0e-14) continue;
    a7[r.slot[t]] += w;
  }
  return sc;
}

uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

}  // namespace

double inlet_profile(double y, double z, double H, double Um) {
  return 16.0 * Um * y * z * (H - y) * (H - z) / std::pow(H, 4.0);  // common.hpp:240-243
}

double seeded_noise(uint64_t seed, int field, size_t idx, double amp) {
  uint64_t key = splitmix64(seed ^ ((uint64_t)(field + 1) << 56));
  uint64_t z = splitmix64(key + (uint64_t)idx);
  double u01 = (double)(z >> 11) * (1.0 / 9007199254740992.0);
  return amp * (2.0 * u01 - 1.0);
}

void SetupData::geometry() {
  const fpc_params &P = par;
  // beta: 10x10 midpoint sub-sampling (immersed_geometry.hpp:1-27)
  int ns = P.nsub_beta;
  std::vector<double> frac((size_t)ns);
  for (int q = 0; q < ns; ++q) frac[(size_t)q] = ((double)q + 0.5) / (double)ns;
  double R2 = P.Rcirc * P.Rcirc;
  beta2d.assign((size_t)nx * ny, 0.0);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double x0 = i * dx, y0 = j * dy;
      int outside = 0;
      for (int a = 0; a < ns; ++a)
        for (int b = 0; b < ns; ++b) {
          double xs = x0 + frac[(size_t)a] * dx, ys = y0 + frac[(size_t)b] * dy;
          if ((xs - P.xcirc) * (xs - P.xcirc) + (ys - P.ycirc) * (ys - P.ycirc) > R2) ++outside;
        }
      beta2d[(size_t)i + (size_t)j * nx] = (double)outside / (double)(ns * ns);
    }
  // open fractions of x-faces and y-faces from circle chords (immersed_geometry.hpp:29-73)
  alpha_u2d.assign((size_t)(nx + 1) * ny, 1.0);
  for (int i = 0; i <= nx; ++i) {
    double x = i * dx;
    if (!(std::abs(x - P.xcirc) < P.Rcirc)) continue;
    double h = std::sqrt(std::max(0.0, R2 - (x - P.xcirc) * (x - P.xcirc)));
    for (int j = 0; j < ny; ++j) {
      double y0 = j * dy, y1 = (j + 1) * dy;
      double ov = std::max(0.0, std::min(y1, P.ycirc + h) - std::max(y0, P.ycirc - h));
      alpha_u2d[(size_t)i + (size_t)j * (nx + 1)] = std::max(0.0, std::min(1.0, 1.0 - ov / dy));
    }
  }
  alpha_v2d.assign((size_t)nx * (ny + 1), 1.0);
  for (int j = 0; j <= ny; ++j) {
    double y = j * dy;
    if (!(std::abs(y - P.ycirc) < P.Rcirc)) continue;
    double h = std::sqrt(std::max(0.0, R2 - (y - P.ycirc) * (y - P.ycirc)));
    for (int i = 0; i < nx; ++i) {
      double x0 = i * dx, x1 = (i + 1) * dx;
      double ov = std::max(0.0, std::min(x1, P.xcirc + h) - std::max(x0, P.xcirc - h));
      alpha_v2d[(size_t)i + (size_t)j * nx] = std::max(0.0, std::min(1.0, 1.0 - ov / dx));
    }
  }
}

void SetupData::masks2d() {
  const fpc_params &P = par;
  double R2 = P.Rcirc * P.Rcirc;
  auto fl = [&](int i, int j) { return fluid2d[(size_t)i + (size_t)j * nx] != 0; };
  au2.assign((size_t)(nx + 1) * ny, 0);
  aqu2.assign((size_t)(nx + 1) * ny, 0);
  av2.assign((size_t)nx * (ny + 1), 0);
  aqv2.assign((size_t)nx * (ny + 1), 0);
  for (int j = 0; j < ny; ++j) {
    double y = (j + 0.5) * dy;
    for (int i = 1; i < nx; ++i) {
      double x = i * dx;
      bool open = ((x - P.xcirc) * (x - P.xcirc) + (y - P.ycirc) * (y - P.ycirc)) > R2 && (fl(i - 1, j) || fl(i, j));
      size_t f = (size_t)i + (size_t)j * (nx + 1);
      au2[f] = open;
      aqu2[f] = open && alpha_u2d[f] > P.alpha_tol;
    }
    aqu2[(size_t)nx + (size_t)j * (nx + 1)] = fl(nx - 1, j);  // outflow face
  }
  for (int j = 1; j < ny; ++j) {
    double y = j * dy;
    for (int i = 0; i < nx; ++i) {
      double x = (i + 0.5) * dx;
      bool open = ((x - P.xcirc) * (x - P.xcirc) + (y - P.ycirc) * (y - P.ycirc)) > R2 && (fl(i, j - 1) || fl(i, j));
      size_t f = (size_t)i + (size_t)j * nx;
      av2[f] = open;
      aqv2[f] = open && alpha_v2d[f] > P.alpha_tol;
    }
  }
}

// Fluid columns with no open projection face are removed until none is left
// (immersed_geometry.hpp:117-160).
void SetupData::prune() {
  fluid2d.assign((size_t)nx * ny, 0);
  for (size_t c = 0; c < fluid2d.size(); ++c) fluid2d[c] = beta2d[c] > par.beta_tol;
  n_pruned = 0;
  for (;;) {
    masks2d();
    std::vector<size_t> drop;
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        size_t c = (size_t)i + (size_t)j * nx;
        if (!fluid2d[c]) continue;
        bool conn = aqu2[(size_t)(i + 1) + (size_t)j * (nx + 1)] || (i > 0 && aqu2[(size_t)i + (size_t)j * (nx + 1)]) ||
                    (j < ny - 1 && aqv2[(size_t)i + (size_t)(j + 1) * nx]) || (j > 0 && aqv2[(size_t)i + (size_t)j * nx]);
        if (!conn) drop.push_back(c);
      }
    if (drop.empty()) break;
    for (size_t c : drop) fluid2d[c] = 0;
    n_pruned += (int)drop.size();
  }
  masks2d();
}

void SetupData::extrude() {
  size_t Nu = (size_t)(nx + 1) * ny * nz, Nv = (size_t)nx * (ny + 1) * nz, Nw = (size_t)nx * ny * (nz + 1),
         Np = (size_t)nx * ny * nz;
  fluid_p.assign(Np, 0);
  active_u.assign(Nu, 0);
  active_qu.assign(Nu, 0);
  active_v.assign(Nv, 0);
  active_qv.assign(Nv, 0);
  active_w.assign(Nw, 0);
  active_qw.assign(Nw, 0);
  for (int k = 0; k < nz; ++k) {
    for (int j = 0; j < ny; ++j) {
      for (int i = 0; i < nx; ++i) fluid_p[pix(i, j, k, nx, ny)] = fluid2d[(size_t)i + (size_t)j * nx];
      for (int i = 0; i <= nx; ++i) {
        active_u[uix(i, j, k, nx, ny)] = au2[(size_t)i + (size_t)j * (nx + 1)];
        active_qu[uix(i, j, k, nx, ny)] = aqu2[(size_t)i + (size_t)j * (nx + 1)];
      }
    }
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i < nx; ++i) {
        active_v[vix(i, j, k, nx, ny)] = av2[(size_t)i + (size_t)j * nx];
        active_qv[vix(i, j, k, nx, ny)] = aqv2[(size_t)i + (size_t)j * nx];
      }
  }
  // w faces (including the two walls) are active next to any fluid cell of the column
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        uint8_t a = fluid2d[(size_t)i + (size_t)j * nx];
        active_w[wix(i, j, k, nx, ny)] = a;
        active_qw[wix(i, j, k, nx, ny)] = a;
      }
  auto number = [](const std::vector<uint8_t> &m, std::vector<int> &id) {
    id.assign(m.size(), 0);
    int c = 0;
    for (size_t t = 0; t < m.size(); ++t)
      if (m[t]) id[t] = ++c;
    return c;
  };
  Nu_ = number(active_u, id_u);
  Nv_ = number(active_v, id_v);
  Nw_ = number(active_w, id_w);
  Np_ = number(fluid_p, id_p);
  Nqu_ = number(active_qu, id_qu);
  Nqv_ = number(active_qv, id_qv);
  Nqw_ = number(active_qw, id_qw);
}

void SetupData::momentum() {
  const fpc_params &P = par;
  double cx = P.nu / (dx * dx), cy = P.nu / (dy * dy), cz = P.nu / (dz * dz), at = P.rho / P.dt;
  auto dist = [&](double a, double b, double c, double d) { return ray_to_circle(a, b, c, d, P); };
  // ---- u: faces i = 1..nx-1 (immersed_operators.hpp:1-62)
  {
    int bx = nx - 1;
    size_t nrow = (size_t)bx * ny * nz;
    Au.assign(7 * nrow, 0.0);
    scale_u_box.assign(nrow, 0.0);
    rhs_u_bc_box.assign(nrow, 0.0);
    scale_u.assign((size_t)Nu_ + 1, 1.0);
    rhs_u_bc.assign((size_t)Nu_ + 1, 0.0);
    auto act = [&](int i, int j, int k) { return active_u[uix(i, j, k, nx, ny)] != 0; };
    auto idf = [&](int i, int j, int k) { return id_u[uix(i, j, k, nx, ny)]; };
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 1; i < nx; ++i) {
          size_t row = ((size_t)k * ny + j) * bx + (i - 1);
          double *a7 = &Au[7 * row];
          if (!act(i, j, k)) { a7[0] = 1.0; continue; }
          double xP = i * dx, yP = (j + 0.5) * dy, zP = (k + 0.5) * dz;
          Row r;
          double dg = at, bc = 0.0;
          if (i > 1 && act(i - 1, j, k)) { r.add(idf(i - 1, j, k), 1, -cx); dg += cx; }
          else if (i == 1) { dg += cx; bc = 0.0 + cx * inlet_profile(yP, zP, P.Ly, 1.0); }
          else dg += cx * (dx / dist(xP, yP, xP - dx, yP));
          if (i < nx - 1 && act(i + 1, j, k)) { r.add(idf(i + 1, j, k), 2, -cx); dg += cx; }
          else if (i != nx - 1) dg += cx * (dx / dist(xP, yP, xP + dx, yP));
          if (j > 0 && act(i, j - 1, k)) { r.add(idf(i, j - 1, k), 3, -cy); dg += cy; }
          else dg += cy * ((j == 0) ? dy / (0.5 * dy) : dy / dist(xP, yP, xP, yP - dy));
          if (j < ny - 1 && act(i, j + 1, k)) { r.add(idf(i, j + 1, k), 4, -cy); dg += cy; }
          else dg += cy * ((j == ny - 1) ? dy / (0.5 * dy) : dy / dist(xP, yP, xP, yP + dy));
          if (k > 0 && act(i, j, k - 1)) { r.add(idf(i, j, k - 1), 5, -cz); dg += cz; }
          else dg += 2.0 * cz;
          if (k < nz - 1 && act(i, j, k + 1)) { r.add(idf(i, j, k + 1), 6, -cz); dg += cz; }
          else dg += 2.0 * cz;
          int id = idf(i, j, k);
          r.add(id, 0, dg);
          double sc = scale_and_embed(r, a7);
          scale_u_box[row] = sc;
          scale_u[(size_t)id] = sc;
          rhs_u_bc_box[row] = bc;
          rhs_u_bc[(size_t)id] = bc;
        }
  }
  // ---- v: faces j = 1..ny-1 (immersed_operators.hpp:64-121)
  {
    int by = ny - 1;
    size_t nrow = (size_t)nx * by * nz;
    Av.assign(7 * nrow, 0.0);
    scale_v_box.assign(nrow, 0.0);
    scale_v.assign((size_t)Nv_ + 1, 1.0);
    rhs_v_bc.assign((size_t)Nv_ + 1, 0.0);
    auto act = [&](int i, int j, int k) { return active_v[vix(i, j, k, nx, ny)] != 0; };
    auto idf = [&](int i, int j, int k) { return id_v[vix(i, j, k, nx, ny)]; };
    for (int k = 0; k < nz; ++k)
      for (int j = 1; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          size_t row = ((size_t)k * by + (j - 1)) * nx + i;
          double *a7 = &Av[7 * row];
          if (!act(i, j, k)) { a7[0] = 1.0; continue; }
          double xP = (i + 0.5) * dx, yP = j * dy;
          Row r;
          double dg = at;
          if (i > 0 && act(i - 1, j, k)) { r.add(idf(i - 1, j, k), 1, -cx); dg += cx; }
          else dg += cx * ((i == 0) ? dx / (0.5 * dx) : dx / dist(xP, yP, xP - dx, yP));
          if (i < nx - 1 && act(i + 1, j, k)) { r.add(idf(i + 1, j, k), 2, -cx); dg += cx; }
          else if (i != nx - 1) dg += cx * (dx / dist(xP, yP, xP + dx, yP));
          if (j > 1 && act(i, j - 1, k)) { r.add(idf(i, j - 1, k), 3, -cy); dg += cy; }
          else dg += cy * ((j == 1) ? dy / dy : dy / dist(xP, yP, xP, yP - dy));
          if (j < ny - 1 && act(i, j + 1, k)) { r.add(idf(i, j + 1, k), 4, -cy); dg += cy; }
          else dg += cy * ((j == ny - 1) ? dy / dy : dy / dist(xP, yP, xP, yP + dy));
          if (k > 0 && act(i, j, k - 1)) { r.add(idf(i, j, k - 1), 5, -cz); dg += cz; }
          else dg += 2.0 * cz;
          if (k < nz - 1 && act(i, j, k + 1)) { r.add(idf(i, j, k + 1), 6, -cz); dg += cz; }
          else dg += 2.0 * cz;
          int id = idf(i, j, k);
          r.add(id, 0, dg);
          double sc = scale_and_embed(r, a7);
          scale_v_box[row] = sc;
          scale_v[(size_t)id] = sc;
        }
  }
  // ---- w: faces k = 1..nz-1 (immersed_operators.hpp:123-177)
  {
    int bz = nz - 1;
    size_t nrow = (size_t)nx * ny * bz;
    Aw.assign(7 * nrow, 0.0);
    scale_w_box.assign(nrow, 0.0);
    scale_w.assign((size_t)Nw_ + 1, 1.0);
    rhs_w_bc.assign((size_t)Nw_ + 1, 0.0);
    auto act = [&](int i, int j, int k) { return active_w[wix(i, j, k, nx, ny)] != 0; };
    auto idf = [&](int i, int j, int k) { return id_w[wix(i, j, k, nx, ny)]; };
    for (int k = 1; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          size_t row = ((size_t)(k - 1) * ny + j) * nx + i;
          double *a7 = &Aw[7 * row];
          if (!act(i, j, k)) { a7[0] = 1.0; continue; }
          double xP = (i + 0.5) * dx, yP = (j + 0.5) * dy;
          Row r;
          double dg = at;
          if (i > 0 && act(i - 1, j, k)) { r.add(idf(i - 1, j, k), 1, -cx); dg += cx; }
          else dg += cx * ((i == 0) ? dx / (0.5 * dx) : dx / dist(xP, yP, xP - dx, yP));
          if (i < nx - 1 && act(i + 1, j, k)) { r.add(idf(i + 1, j, k), 2, -cx); dg += cx; }
          else if (i != nx - 1) dg += cx * (dx / dist(xP, yP, xP + dx, yP));
          if (j > 0 && act(i, j - 1, k)) { r.add(idf(i, j - 1, k), 3, -cy); dg += cy; }
          else dg += cy * ((j == 0) ? dy / (0.5 * dy) : dy / dist(xP, yP, xP, yP - dy));
          if (j < ny - 1 && act(i, j + 1, k)) { r.add(idf(i, j + 1, k), 4, -cy); dg += cy; }
          else dg += cy * ((j == ny - 1) ? dy / (0.5 * dy) : dy / dist(xP, yP, xP, yP + dy));
          if (k > 1 && act(i, j, k - 1)) { r.add(idf(i, j, k - 1), 5, -cz); dg += cz; }
          else dg += cz;
          if (k < nz - 1 && act(i, j, k + 1)) { r.add(idf(i, j, k + 1), 6, -cz); dg += cz; }
          else dg += cz;
          int id = idf(i, j, k);
          r.add(id, 0, dg);
          double sc = scale_and_embed(r, a7);
          scale_w_box[row] = sc;
          scale_w[(size_t)id] = sc;
        }
  }
}

// D, G and Ap = -(D G) (immersed_operators.hpp:179-287, main.cu:101-103).
void SetupData::projection() {
  const fpc_params &P = par;
  int nQu = Nqu_, nQv = Nqv_;
  dcoef.assign((size_t)nx * ny * 6, 0.0);
  div_bc_plane.assign((size_t)ny * nz, 0.0);
  div_bc.assign((size_t)Np_ + 1, 0.0);
  Ap.assign(7 * (size_t)nx * ny * nz, 0.0);
  // divergence coefficients per (i, j) column: z-invariant
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      if (!fluid2d[(size_t)i + (size_t)j * nx]) continue;
      double beta = beta2d[(size_t)i + (size_t)j * nx];
      double *c = &dcoef[6 * ((size_t)j * nx + i)];
      size_t fe = (size_t)(i + 1) + (size_t)j * (nx + 1), fw = (size_t)i + (size_t)j * (nx + 1);
      size_t fn = (size_t)i + (size_t)(j + 1) * nx, fs = (size_t)i + (size_t)j * nx;
      if (i >= 1 && aqu2[fw]) c[0] = -alpha_u2d[fw] / (beta * dx);
      if (aqu2[fe]) c[1] = alpha_u2d[fe] / (beta * dx);
      if (j >= 1 && aqv2[fs]) c[2] = -alpha_v2d[fs] / (beta * dy);
      if (j <= ny - 2 && aqv2[fn]) c[3] = alpha_v2d[fn] / (beta * dy);
      c[4] = -1.0 / (beta * dz);
      c[5] = 1.0 / (beta * dz);
      for (int q = 0; q < 6; ++q)
        if (std::abs(c[q]) < 1.0e-30 || std::abs(c[q]) <= 1.0e-14) c[q] = 0.0;  // add_entry / compress_zeros
    }
  const double gx = 1.0 / dx, gy = 1.0 / dy, gz = 1.0 / dz, gout = 2.0 / dx;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t cell = pix(i, j, k, nx, ny);
        double *a7 = &Ap[7 * cell];
        if (!fluid_p[cell]) { a7[0] = 1.0; continue; }
        int row = id_p[cell];
        if (i == 0) {
          double beta = beta2d[(size_t)j * nx];
          div_bc[(size_t)row] = 0.0 - inlet_profile((j + 0.5) * dy, (k + 0.5) * dz, P.Ly, P.Um) / (beta * dx);
          div_bc_plane[(size_t)j + (size_t)ny * k] = div_bc[(size_t)row];
        }
        const double *c = &dcoef[6 * ((size_t)j * nx + i)];
        // D row in insertion order: east, west, north, south, top, bottom.
        int key[6], n = 0;
        double tdiag[6];   // c_f * g_f(P): this cell's term of Ap0's diagonal
        int nslot[6];      // slot of the neighbour reached through the face, -1 if none
        double toff[6];    // c_f * g_f(neighbour)
        auto face = [&](int kk, double cf, double gP, int slot, bool nbr_fluid, double gN) {
          if (cf == 0.0) return;
          key[n] = kk;
          tdiag[n] = cf * gP;
          nslot[n] = nbr_fluid ? slot : -1;
          toff[n] = nbr_fluid ? cf * gN : 0.0;
          ++n;
        };
        if (c[1] != 0.0) {
          if (i + 1 < nx) face(id_qu[uix(i + 1, j, k, nx, ny)], c[1], -gx, 2, fluid_p[pix(i + 1, j, k, nx, ny)] != 0, gx);
          else face(id_qu[uix(nx, j, k, nx, ny)], c[1], -gout, 2, false, 0.0);
        }
        if (c[0] != 0.0) face(id_qu[uix(i, j, k, nx, ny)], c[0], gx, 1, fluid_p[pix(i - 1, j, k, nx, ny)] != 0, -gx);
        if (c[3] != 0.0) face(nQu + id_qv[vix(i, j + 1, k, nx, ny)], c[3], -gy, 4, fluid_p[pix(i, j + 1, k, nx, ny)] != 0, gy);
        if (c[2] != 0.0) face(nQu + id_qv[vix(i, j, k, nx, ny)], c[2], gy, 3, fluid_p[pix(i, j - 1, k, nx, ny)] != 0, -gy);
        if (k < nz - 1 && c[5] != 0.0)
          face(nQu + nQv + id_qw[wix(i, j, k + 1, nx, ny)], c[5], -gz, 6, fluid_p[pix(i, j, k + 1, nx, ny)] != 0, gz);
        if (k > 0 && c[4] != 0.0)
          face(nQu + nQv + id_qw[wix(i, j, k, nx, ny)], c[4], gz, 5, fluid_p[pix(i, j, k - 1, nx, ny)] != 0, -gz);
        int ord[6];
        map_order(key, n, ord);
        double dsum = 0.0;
        bool any = false;
        for (int t = 0; t < n; ++t) {
          double v = tdiag[ord[t]];
          if (std::abs(v) < 1.0e-30) continue;  // add_entry in the product
          dsum = any ? dsum + v : 0.0 + v;
          any = true;
        }
        if (any && !(std::abs(dsum) <= 1.0e-14)) a7[0] = 0.0 + (-dsum);
        for (int t = 0; t < n; ++t) {
          if (nslot[t] < 0) continue;
          double v = toff[t];
          if (std::abs(v) < 1.0e-30 || std::abs(v) <= 1.0e-14) continue;
          a7[nslot[t]] = 0.0 + (-v);
        }
      }
}

// CSR forms of D and G with sorted columns (gpu_dense_ops.hpp:32-48), for the
// 1:1 kernel API.
void SetupData::csr() {
  int nQu = Nqu_, nQv = Nqv_;
  D_rowptr.assign((size_t)Np_ + 1, 0);
  D_colind.clear();
  D_vals.clear();
  int r = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        if (!fluid_p[pix(i, j, k, nx, ny)]) continue;
        const double *c = &dcoef[6 * ((size_t)j * nx + i)];
        auto put = [&](int col, double v) { D_colind.push_back(col); D_vals.push_back(v); };
        if (c[0] != 0.0) put(id_qu[uix(i, j, k, nx, ny)], c[0]);
        if (c[1] != 0.0) put(id_qu[uix(i + 1, j, k, nx, ny)], c[1]);
        if (c[2] != 0.0) put(nQu + id_qv[vix(i, j, k, nx, ny)], c[2]);
        if (c[3] != 0.0) put(nQu + id_qv[vix(i, j + 1, k, nx, ny)], c[3]);
        if (k > 0 && c[4] != 0.0) put(nQu + nQv + id_qw[wix(i, j, k, nx, ny)], c[4]);
        if (k < nz - 1 && c[5] != 0.0) put(nQu + nQv + id_qw[wix(i, j, k + 1, nx, ny)], c[5]);
        D_rowptr[(size_t)++r] = (int)D_colind.size();
      }
  int Nq = Nqu_ + Nqv_ + Nqw_;
  G_rowptr.assign((size_t)Nq + 1, 0);
  G_colind.clear();
  G_vals.clear();
  r = 0;
  auto put2 = [&](int pl, double vl, int pr, double vr) {
    if (pl > 0) { G_colind.push_back(pl); G_vals.push_back(vl); }
    if (pr > 0) { G_colind.push_back(pr); G_vals.push_back(vr); }
    G_rowptr[(size_t)++r] = (int)G_colind.size();
  };
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        if (!active_qu[uix(i, j, k, nx, ny)]) continue;
        if (i == nx) put2(id_p[pix(nx - 1, j, k, nx, ny)], -2.0 / dx, 0, 0.0);
        else put2(id_p[pix(i - 1, j, k, nx, ny)], -1.0 / dx, id_p[pix(i, j, k, nx, ny)], 1.0 / dx);
      }
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i < nx; ++i) {
        if (!active_qv[vix(i, j, k, nx, ny)]) continue;
        put2(id_p[pix(i, j - 1, k, nx, ny)], -1.0 / dy, id_p[pix(i, j, k, nx, ny)], 1.0 / dy);
      }
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        if (!active_qw[wix(i, j, k, nx, ny)]) continue;
        if (k == 0 || k == nz) put2(0, 0.0, 0, 0.0);  // wall faces carry ids but no gradient
        else put2(id_p[pix(i, j, k - 1, nx, ny)], -1.0 / dz, id_p[pix(i, j, k, nx, ny)], 1.0 / dz);
      }
}

// Drag/lift functional weights (no reference counterpart):
//  pressure: F_x += p (alpha_west - alpha_east) dy dz, F_y += p (alpha_south - alpha_north) dx dz;
//  viscous: every u (v) closure against the cylinder adds rho nu A u_P / d_P to F_x (F_y).
void SetupData::forces() {
  const fpc_params &P = par;
  double rn = P.rho * P.nu;
  auto dist = [&](double a, double b, double c, double d) { return ray_to_circle(a, b, c, d, P); };
  auto add = [&](std::vector<long long> &idx, std::vector<double> &wt, size_t t, double area, double dP) {
    if (std::isfinite(dP) && dP > 0.0) {
      idx.push_back((long long)t);
      wt.push_back(rn * area / dP);
    }
  };
  fu_idx.clear(); fu_w.clear(); fv_idx.clear(); fv_w.clear(); px_idx.clear(); px_w.clear(); py_idx.clear(); py_w.clear();
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 1; i < nx; ++i) {
        size_t t = uix(i, j, k, nx, ny);
        if (!active_u[t]) continue;
        double xP = i * dx, yP = (j + 0.5) * dy;
        if (i > 1 && !active_u[uix(i - 1, j, k, nx, ny)]) add(fu_idx, fu_w, t, dy * dz, dist(xP, yP, xP - dx, yP));
        if (i < nx - 1 && !active_u[uix(i + 1, j, k, nx, ny)]) add(fu_idx, fu_w, t, dy * dz, dist(xP, yP, xP + dx, yP));
        if (j > 0 && !active_u[uix(i, j - 1, k, nx, ny)]) add(fu_idx, fu_w, t, dx * dz, dist(xP, yP, xP, yP - dy));
        if (j < ny - 1 && !active_u[uix(i, j + 1, k, nx, ny)]) add(fu_idx, fu_w, t, dx * dz, dist(xP, yP, xP, yP + dy));
      }
  for (int k = 0; k < nz; ++k)
    for (int j = 1; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t t = vix(i, j, k, nx, ny);
        if (!active_v[t]) continue;
        double xP = (i + 0.5) * dx, yP = j * dy;
        if (i > 0 && !active_v[vix(i - 1, j, k, nx, ny)]) add(fv_idx, fv_w, t, dy * dz, dist(xP, yP, xP - dx, yP));
        if (i < nx - 1 && !active_v[vix(i + 1, j, k, nx, ny)]) add(fv_idx, fv_w, t, dy * dz, dist(xP, yP, xP + dx, yP));
        if (j > 1 && !active_v[vix(i, j - 1, k, nx, ny)]) add(fv_idx, fv_w, t, dx * dz, dist(xP, yP, xP, yP - dy));
        if (j < ny - 1 && !active_v[vix(i, j + 1, k, nx, ny)]) add(fv_idx, fv_w, t, dx * dz, dist(xP, yP, xP, yP + dy));
      }
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t c = pix(i, j, k, nx, ny);
        if (!fluid_p[c]) continue;
        double aw = alpha_u2d[(size_t)i + (size_t)j * (nx + 1)], ae = alpha_u2d[(size_t)(i + 1) + (size_t)j * (nx + 1)];
        double as = alpha_v2d[(size_t)i + (size_t)j * nx], an = alpha_v2d[(size_t)i + (size_t)(j + 1) * nx];
        if (aw != ae) { px_idx.push_back((long long)c); px_w.push_back((aw - ae) * dy * dz); }
        if (as != an) { py_idx.push_back((long long)c); py_w.push_back((as - an) * dx * dz); }
      }
}

void SetupData::build(const fpc_params &p) {
  par = p;
  nx = p.nx; ny = p.ny; nz = p.nz;
  dx = p.Lx / nx; dy = p.Ly / ny; dz = p.Lz / nz;
  geometry();
  prune();
  extrude();
  momentum();
  projection();
  csr();
  forces();
  uin.assign((size_t)ny * nz, 0.0);  // immersed_operators.hpp:921-925
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) uin[(size_t)j + (size_t)ny * k] = inlet_profile((j + 0.5) * dy, (k + 0.5) * dz, p.Ly, p.Um);
}

// Initial fields: zero with the inlet face (main.cu:123-126) or, with
// init_profile, the inlet profile on active u faces; seeded noise on active
// faces when noise_amp > 0; then enforce_bc_and_cleanup (immersed_operators.hpp:1036-1053).
void SetupData::initial_fields(std::vector<double> &u, std::vector<double> &v, std::vector<double> &w,
                               std::vector<double> &p) const {
  u.assign((size_t)(nx + 1) * ny * nz, 0.0);
  v.assign((size_t)nx * (ny + 1) * nz, 0.0);
  w.assign((size_t)nx * ny * (nz + 1), 0.0);
  p.assign((size_t)nx * ny * nz, 0.0);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 1; i < nx; ++i) {
        size_t t = uix(i, j, k, nx, ny);
        if (!active_u[t]) continue;
        double val = par.init_profile ? inlet_profile((j + 0.5) * dy, (k + 0.5) * dz, par.Ly, par.Um) : 0.0;
        if (par.noise_amp > 0.0) val += seeded_noise(par.seed, 0, t, par.noise_amp);
        u[t] = val;
      }
  if (par.noise_amp > 0.0) {
    for (int k = 0; k < nz; ++k)
      for (int j = 1; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          size_t t = vix(i, j, k, nx, ny);
          if (active_v[t]) v[t] = seeded_noise(par.seed, 1, t, par.noise_amp);
        }
    for (int k = 1; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          size_t t = wix(i, j, k, nx, ny);
          if (active_w[t]) w[t] = seeded_noise(par.seed, 2, t, par.noise_amp);
        }
  }
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) {
      u[uix(0, j, k, nx, ny)] = uin[(size_t)j + (size_t)ny * k];
      u[uix(nx, j, k, nx, ny)] = u[uix(nx - 1, j, k, nx, ny)];
    }
  for (int k = 0; k < nz; ++k)
    for (int i = 0; i < nx; ++i) {
      v[vix(i, 0, k, nx, ny)] = 0.0;
      v[vix(i, ny, k, nx, ny)] = 0.0;
    }
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      w[wix(i, j, 0, nx, ny)] = 0.0;
      w[wix(i, j, nz, nx, ny)] = 0.0;
    }
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 1; i < nx; ++i)
        if (!active_u[uix(i, j, k, nx, ny)]) u[uix(i, j, k, nx, ny)] = 0.0;
  for (int k = 0; k < nz; ++k)
    for (int j = 1; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        if (!active_v[vix(i, j, k, nx, ny)]) v[vix(i, j, k, nx, ny)] = 0.0;
  for (int k = 1; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        if (!active_w[wix(i, j, k, nx, ny)]) w[wix(i, j, k, nx, ny)] = 0.0;
}

namespace {
struct ViewC {
  const void *p = nullptr;
  long long n = -1;
  int es = 0;
};
template <class T>
ViewC vw(const std::vector<T> &v) {
  ViewC r;
  r.p = v.data();
  r.n = (long long)v.size();
  r.es = (int)sizeof(T);
  return r;
}
}  // namespace

int SetupData::view(const char *cname, const void **p, long long *n, int *es) const {
  std::string s(cname);
  ViewC r;
  if (s == "beta2d") r = vw(beta2d);
  else if (s == "alpha_u2d") r = vw(alpha_u2d);
  else if (s == "alpha_v2d") r = vw(alpha_v2d);
  else if (s == "fluid2d") r = vw(fluid2d);
  else if (s == "fluid_p") r = vw(fluid_p);
  else if (s == "active_u") r = vw(active_u);
  else if (s == "active_v") r = vw(active_v);
  else if (s == "active_w") r = vw(active_w);
  else if (s == "active_qu") r = vw(active_qu);
  else if (s == "active_qv") r = vw(active_qv);
  else if (s == "active_qw") r = vw(active_qw);
  else if (s == "id_u") r = vw(id_u);
  else if (s == "id_v") r = vw(id_v);
  else if (s == "id_w") r = vw(id_w);
  else if (s == "id_p") r = vw(id_p);
  else if (s == "id_qu") r = vw(id_qu);
  else if (s == "id_qv") r = vw(id_qv);
  else if (s == "id_qw") r = vw(id_qw);
  else if (s == "rhs_u_bc") r = vw(rhs_u_bc);
  else if (s == "rhs_v_bc") r = vw(rhs_v_bc);
  else if (s == "rhs_w_bc") r = vw(rhs_w_bc);
  else if (s == "div_bc") r = vw(div_bc);
  else if (s == "scale_u") r = vw(scale_u);
  else if (s == "scale_v") r = vw(scale_v);
  else if (s == "scale_w") r = vw(scale_w);
  else if (s == "Au_vals") r = vw(Au);
  else if (s == "Av_vals") r = vw(Av);
  else if (s == "Aw_vals") r = vw(Aw);
  else if (s == "Ap_vals") r = vw(Ap);
  else if (s == "D_rowptr") r = vw(D_rowptr);
  else if (s == "D_colind") r = vw(D_colind);
  else if (s == "D_vals") r = vw(D_vals);
  else if (s == "G_rowptr") r = vw(G_rowptr);
  else if (s == "G_colind") r = vw(G_colind);
  else if (s == "G_vals") r = vw(G_vals);
  else if (s == "uin") r = vw(uin);
  else if (s == "scale_u_box") r = vw(scale_u_box);
  else if (s == "scale_v_box") r = vw(scale_v_box);
  else if (s == "scale_w_box") r = vw(scale_w_box);
  else if (s == "rhs_u_bc_box") r = vw(rhs_u_bc_box);
  else if (s == "dcoef") r = vw(dcoef);
  else if (s == "div_bc_plane") r = vw(div_bc_plane);
  else return -1;
  *p = r.p;
  *n = r.n;
  *es = r.es;
  return 0;
}

long long SetupData::count(const char *cname) const {
  std::string s(cname);
  if (s == "Nu") return Nu_;
  if (s == "Nv") return Nv_;
  if (s == "Nw") return Nw_;
  if (s == "Np") return Np_;
  if (s == "Nqu") return Nqu_;
  if (s == "Nqv") return Nqv_;
  if (s == "Nqw") return Nqw_;
  if (s == "n_pruned") return n_pruned;
  return -1;
}

}  // namespace host
}  // namespace fpc

// ---------------------------------------------------------------- C ABI
struct fpc_setup_s {
  fpc::host::SetupData d;
};

extern "C" {

void fpc_default_params(fpc_params *p) {
  // common.hpp:45-85
  std::memset(p, 0, sizeof(*p));
  p->nx = 32; p->ny = 12; p->nz = 12;
  p->Lx = 2.2; p->Ly = 0.41; p->Lz = 0.41;
  p->rho = 1.0; p->nu = 1.0e-3; p->Um = 0.45; p->dt = 2.0e-3;
  p->nsteps = 10; p->print_every = 1; p->inlet_ramp_steps = 200; p->use_zero_initial_guess = 1;
  p->nsub_beta = 10; p->beta_tol = 1.0e-8; p->alpha_tol = 1.0e-10;
  p->xcirc = 0.2; p->ycirc = 0.2; p->Rcirc = 0.05;
  p->vel_maxit = 250; p->vel_kdim = 50; p->vel_tol = 1.0e-6; p->vel_jacobi_iters = 1;
  p->p_maxit = 250; p->p_kdim = 50; p->p_tol = 1.0e-9;
  p->p_pc_maxit = 1; p->p_pc_relax_type = 1; p->p_pc_rap_type = 1;
  p->init_profile = 0; p->noise_amp = 0.0; p->seed = 0; p->mg_precision = 64;
  p->vel_print = 0; p->p_print = 0;
  p->write_vtk = 1; p->vtk_every = 0; p->device = 0;
  std::snprintf(p->vtk_filename, sizeof(p->vtk_filename), "%s", "st3d_channel_cylinder_struct_mac_cuda.vtk");
}

int fpc_parse_args(int argc, char **argv, fpc_params *p) {
  // common.hpp:87-127 flag names; atoi/atof semantics, unknown flags ignored.
  for (int a = 1; a < argc; ++a) {
    const char *f = argv[a];
    if (a + 1 >= argc) break;
    const char *v = argv[a + 1];
    bool used = true;
    if (!std::strcmp(f, "-nx")) p->nx = std::atoi(v);
    else if (!std::strcmp(f, "-ny")) p->ny = std::atoi(v);
    else if (!std::strcmp(f, "-nz")) p->nz = std::atoi(v);
    else if (!std::strcmp(f, "-Lx")) p->Lx = std::atof(v);
    else if (!std::strcmp(f, "-Ly")) p->Ly = std::atof(v);
    else if (!std::strcmp(f, "-Lz")) p->Lz = std::atof(v);
    else if (!std::strcmp(f, "-rho")) p->rho = std::atof(v);
    else if (!std::strcmp(f, "-nu")) p->nu = std::atof(v);
    else if (!std::strcmp(f, "-Um")) p->Um = std::atof(v);
    else if (!std::strcmp(f, "-dt")) p->dt = std::atof(v);
    else if (!std::strcmp(f, "-nsteps")) p->nsteps = std::atoi(v);
    else if (!std::strcmp(f, "-print-every")) p->print_every = std::atoi(v);
    else if (!std::strcmp(f, "-inlet-ramp-steps")) p->inlet_ramp_steps = std::atoi(v);
    else if (!std::strcmp(f, "-use-zero-initial-guess")) p->use_zero_initial_guess = std::atoi(v);
    else if (!std::strcmp(f, "-nsub-beta")) p->nsub_beta = std::atoi(v);
    else if (!std::strcmp(f, "-beta-tol")) p->beta_tol = std::atof(v);
    else if (!std::strcmp(f, "-alpha-tol")) p->alpha_tol = std::atof(v);
    else if (!std::strcmp(f, "-xcirc")) p->xcirc = std::atof(v);
    else if (!std::strcmp(f, "-ycirc")) p->ycirc = std::atof(v);
    else if (!std::strcmp(f, "-Rcirc")) p->Rcirc = std::atof(v);
    else if (!std::strcmp(f, "-vel-maxit")) p->vel_maxit = std::atoi(v);
    else if (!std::strcmp(f, "-vel-kdim")) p->vel_kdim = std::atoi(v);
    else if (!std::strcmp(f, "-vel-tol")) p->vel_tol = std::atof(v);
    else if (!std::strcmp(f, "-vel-print")) p->vel_print = std::atoi(v);
    else if (!std::strcmp(f, "-vel-jacobi-iters")) p->vel_jacobi_iters = std::atoi(v);
    else if (!std::strcmp(f, "-p-maxit")) p->p_maxit = std::atoi(v);
    else if (!std::strcmp(f, "-p-kdim")) p->p_kdim = std::atoi(v);
    else if (!std::strcmp(f, "-p-tol")) p->p_tol = std::atof(v);
    else if (!std::strcmp(f, "-p-print")) p->p_print = std::atoi(v);
    else if (!std::strcmp(f, "-p-pc-maxit")) p->p_pc_maxit = std::atoi(v);
    else if (!std::strcmp(f, "-p-pc-relax-type")) p->p_pc_relax_type = std::atoi(v);
    else if (!std::strcmp(f, "-p-pc-rap-type")) p->p_pc_rap_type = std::atoi(v);
    else if (!std::strcmp(f, "-init-profile")) p->init_profile = std::atoi(v);
    else if (!std::strcmp(f, "-noise-amp")) p->noise_amp = std::atof(v);
    else if (!std::strcmp(f, "-seed")) p->seed = std::strtoull(v, nullptr, 10);
    else if (!std::strcmp(f, "-mg-precision")) p->mg_precision = std::atoi(v);
    else if (!std::strcmp(f, "-write-vtk")) p->write_vtk = std::atoi(v);
    else if (!std::strcmp(f, "-vtk-every")) p->vtk_every = std::atoi(v);
    else if (!std::strcmp(f, "-device")) p->device = std::atoi(v);
    else if (!std::strcmp(f, "-vtk")) std::snprintf(p->vtk_filename, sizeof(p->vtk_filename), "%s", v);
    else used = false;
    if (used) ++a;
  }
  return 0;
}

int fpc_setup_create(const fpc_params *par, fpc_setup *out) {
  if (!par || !out) return FPC_ERROR_ARG;
  if (par->nx < 2 || par->ny < 2 || par->nz < 2) return FPC_ERROR_ARG;
  fpc_setup s = new fpc_setup_s();
  s->d.build(*par);
  *out = s;
  return 0;
}

long long fpc_setup_len(fpc_setup s, const char *name) {
  const void *p;
  long long n;
  int es;
  if (!s || s->d.view(name, &p, &n, &es)) return -1;
  return n;
}

int fpc_setup_elsize(fpc_setup s, const char *name) {
  const void *p;
  long long n;
  int es;
  if (!s || s->d.view(name, &p, &n, &es)) return -1;
  return es;
}

int fpc_setup_copy(fpc_setup s, const char *name, void *dst, long long nbytes) {
  const void *p;
  long long n;
  int es;
  if (!s || s->d.view(name, &p, &n, &es)) return FPC_ERROR_ARG;
  if (nbytes < n * es) return FPC_ERROR_ARG;
  std::memcpy(dst, p, (size_t)(n * es));
  return 0;
}

long long fpc_setup_int(fpc_setup s, const char *name) { return s ? s->d.count(name) : -1; }

int fpc_setup_destroy(fpc_setup s) {
  delete s;
  return 0;
}

}  // extern "C"

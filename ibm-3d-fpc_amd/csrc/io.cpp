// io.cpp -- on-disk formats of the channel solver (libfpc, host code).
//
//  * The cell-centred ASCII VTK file of the reference driver
//    (write_vtk_cellcentered, immersed_operators.hpp:1055-1076): same header
//    lines, same scalars in the same order (ux, uy, uz as face averages,
//    pressure, fluid_mask), every double as "%.16e", so the files are byte
//    for byte those the reference writes for the same fields.  Numbers are
//    formatted with std::to_chars (scientific, 16 digits), which gives the
//    same characters as printf's "%.16e" (exponent of at least two digits,
//    correctly rounded) at a fraction of the cost, and written in large blocks.
//  * The per-step file name: base name with "_%06d" before the extension
//    (vtk_filename_with_step, solver_helpers.hpp:470-482).
//  * A raw dump of the staggered fields (no reference counterpart): a 40-byte
//    header ("FPCFLD01", int32 nx, ny, nz, step, int64 total doubles, int32 0)
//    then u, v, w, p as little-endian float64 in the full staggered layouts.
#include <charconv>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fpc_sim.h"
#include "fpc_struct.h"

namespace {

// Appends values as "%.16e\n" lines into a buffer flushed to fp in large blocks.
class LineWriter {
 public:
  explicit LineWriter(std::FILE *fp) : fp_(fp) { buf_.reserve(kBlock + 64); }
  ~LineWriter() { flush(); }
  void text(const char *s) {
    buf_.append(s);
    if (buf_.size() >= kBlock) flush();
  }
  void real(double x) {
    char tmp[40];
    auto r = std::to_chars(tmp, tmp + sizeof(tmp) - 1, x, std::chars_format::scientific, 16);
    *r.ptr++ = '\n';
    buf_.append(tmp, (size_t)(r.ptr - tmp));
    if (buf_.size() >= kBlock) flush();
  }
  void flush() {
    if (!buf_.empty()) {
      if (std::fwrite(buf_.data(), 1, buf_.size(), fp_) != buf_.size()) ok_ = false;
      buf_.clear();
    }
  }
  bool ok() const { return ok_; }

 private:
  static constexpr size_t kBlock = 1 << 22;
  std::FILE *fp_;
  std::string buf_;
  bool ok_ = true;
};

inline size_t iu(int i, int j, int k, int nx, int ny) { return ((size_t)k * ny + j) * (size_t)(nx + 1) + i; }
inline size_t iv(int i, int j, int k, int nx, int ny) { return ((size_t)k * (ny + 1) + j) * (size_t)nx + i; }
inline size_t iw(int i, int j, int k, int nx, int ny) { return ((size_t)k * ny + j) * (size_t)nx + i; }

}  // namespace

extern "C" {

int fpc_write_vtk_cellcentered(const char *filename, const double *u, const double *v, const double *w,
                               const double *p, const unsigned char *fluid_p, int nx, int ny, int nz, double Lx,
                               double Ly, double Lz) {
  if (!filename || !u || !v || !w || !p || !fluid_p || nx < 1 || ny < 1 || nz < 1) return FPC_ERROR_ARG;
  std::FILE *fp = std::fopen(filename, "w");
  if (!fp) return FPC_ERROR_GENERIC;
  const double dx = Lx / nx, dy = Ly / ny, dz = Lz / nz;
  bool ok = true;
  {
    LineWriter out(fp);
    char line[256];
    out.text("# vtk DataFile Version 3.0\n");
    out.text("3D immersed cylinder MAC cell-centered\nASCII\nDATASET STRUCTURED_POINTS\n");
    std::snprintf(line, sizeof(line), "DIMENSIONS %d %d %d\n", nx + 1, ny + 1, nz + 1);
    out.text(line);
    std::snprintf(line, sizeof(line), "ORIGIN 0 0 0\nSPACING %.16e %.16e %.16e\n", dx, dy, dz);
    out.text(line);
    std::snprintf(line, sizeof(line), "CELL_DATA %d\n", nx * ny * nz);
    out.text(line);
    out.text("SCALARS ux double 1\nLOOKUP_TABLE default\n");
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) out.real(0.5 * (u[iu(i, j, k, nx, ny)] + u[iu(i + 1, j, k, nx, ny)]));
    out.text("SCALARS uy double 1\nLOOKUP_TABLE default\n");
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) out.real(0.5 * (v[iv(i, j, k, nx, ny)] + v[iv(i, j + 1, k, nx, ny)]));
    out.text("SCALARS uz double 1\nLOOKUP_TABLE default\n");
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) out.real(0.5 * (w[iw(i, j, k, nx, ny)] + w[iw(i, j, k + 1, nx, ny)]));
    out.text("SCALARS pressure double 1\nLOOKUP_TABLE default\n");
    const size_t N = (size_t)nx * ny * nz;
    for (size_t c = 0; c < N; ++c) out.real(p[c]);
    out.text("SCALARS fluid_mask int 1\nLOOKUP_TABLE default\n");
    for (size_t c = 0; c < N; ++c) out.text(fluid_p[c] ? "1\n" : "0\n");
    out.flush();
    ok = out.ok();
  }
  if (std::fclose(fp) != 0) ok = false;
  return ok ? 0 : FPC_ERROR_GENERIC;
}

int fpc_vtk_filename_with_step(const char *base, int step, char *out, int outlen) {
  if (!base || !out || outlen < 1) return FPC_ERROR_ARG;
  std::string b(base), stem = b, ext;
  const size_t dot = b.find_last_of('.');
  if (dot != std::string::npos) {
    stem = b.substr(0, dot);
    ext = b.substr(dot);
  }
  char num[64];
  std::snprintf(num, sizeof(num), "_%06d", step);
  const std::string name = stem + num + ext;
  if ((int)name.size() + 1 > outlen) return FPC_ERROR_ARG;
  std::memcpy(out, name.c_str(), name.size() + 1);
  return 0;
}

int fpc_write_fields(const char *path, int step, int nx, int ny, int nz, const double *u, const double *v,
                     const double *w, const double *p) {
  if (!path || !u || !v || !w || !p || nx < 1 || ny < 1 || nz < 1) return FPC_ERROR_ARG;
  const size_t Nu = (size_t)(nx + 1) * ny * nz, Nv = (size_t)nx * (ny + 1) * nz, Nw = (size_t)nx * ny * (nz + 1),
               Np = (size_t)nx * ny * nz;
  std::FILE *fp = std::fopen(path, "wb");
  if (!fp) return FPC_ERROR_GENERIC;
  unsigned char head[40] = {};
  std::memcpy(head, "FPCFLD01", 8);
  const int32_t dims[4] = {nx, ny, nz, step};
  std::memcpy(head + 8, dims, sizeof(dims));
  const int64_t total = (int64_t)(Nu + Nv + Nw + Np);
  std::memcpy(head + 24, &total, sizeof(total));
  bool ok = std::fwrite(head, 1, sizeof(head), fp) == sizeof(head);
  ok = ok && std::fwrite(u, sizeof(double), Nu, fp) == Nu;
  ok = ok && std::fwrite(v, sizeof(double), Nv, fp) == Nv;
  ok = ok && std::fwrite(w, sizeof(double), Nw, fp) == Nw;
  ok = ok && std::fwrite(p, sizeof(double), Np, fp) == Np;
  if (std::fclose(fp) != 0) ok = false;
  return ok ? 0 : FPC_ERROR_GENERIC;
}

}  // extern "C"

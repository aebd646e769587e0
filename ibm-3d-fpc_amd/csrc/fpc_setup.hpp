// fpc_setup.hpp -- host-side setup data of one channel (internal to libfpc).
#pragma once

#include <cstdint>
#include <vector>

#include "fpc_dense.h"
#include "fpc_sim.h"
#include "fpc_struct.h"

namespace fpc {
namespace host {

double inlet_profile(double y, double z, double H, double Um);
double seeded_noise(uint64_t seed, int field, size_t idx, double amp);

struct SetupData {
  fpc_params par{};
  int nx = 0, ny = 0, nz = 0;
  double dx = 0, dy = 0, dz = 0;
  std::vector<double> beta2d, alpha_u2d, alpha_v2d;
  std::vector<uint8_t> fluid2d, au2, av2, aqu2, aqv2;
  int n_pruned = 0;
  std::vector<uint8_t> fluid_p, active_u, active_v, active_w, active_qu, active_qv, active_qw;
  std::vector<int> id_u, id_v, id_w, id_p, id_qu, id_qv, id_qw;
  int Nu_ = 0, Nv_ = 0, Nw_ = 0, Np_ = 0, Nqu_ = 0, Nqv_ = 0, Nqw_ = 0;
  std::vector<double> Au, Av, Aw, Ap;  // AoS 7-point boxes
  std::vector<double> scale_u, scale_v, scale_w, rhs_u_bc, rhs_v_bc, rhs_w_bc, div_bc;  // 1-based
  std::vector<double> scale_u_box, scale_v_box, scale_w_box, rhs_u_bc_box;               // per box row
  std::vector<double> dcoef, div_bc_plane, uin;
  std::vector<int> D_rowptr, D_colind, G_rowptr, G_colind;
  std::vector<double> D_vals, G_vals;
  std::vector<long long> fu_idx, fv_idx, px_idx, py_idx;
  std::vector<double> fu_w, fv_w, px_w, py_w;

  void build(const fpc_params &p);
  void initial_fields(std::vector<double> &u, std::vector<double> &v, std::vector<double> &w,
                      std::vector<double> &p) const;
  int view(const char *name, const void **ptr, long long *n, int *elsize) const;
  long long count(const char *name) const;

 private:
  void geometry();
  void masks2d();
  void prune();
  void extrude();
  void momentum();
  void projection();
  void csr();
  void forces();
};

}  // namespace host

// Asynchronous fused stages (dense_ops.hip).
int stage_explicit_async(const fpc_flow_ops *o, double *u, double *v, double *w, const double *p, const double *uin,
                         double *ru, double *xu, double *rv, double *xv, double *rw, double *xw, int zero);
int stage_scatter_async(const fpc_flow_ops *o, const double *xu, const double *xv, const double *xw, const double *u,
                        const double *v, const double *w, double *ut, double *vt, double *wt, const double *uin);
int stage_divergence_async(const fpc_flow_ops *o, const double *ut, const double *vt, const double *wt, double ramp,
                           double *rhs, double *x0);
int stage_projection_async(const fpc_flow_ops *o, const double *phi, const double *ut, const double *vt,
                           const double *wt, double *u, double *v, double *w, double *p, const double *uin,
                           double *qout);
int stage_diagnostics_async(const fpc_flow_ops *o, const double *u, const double *v, const double *w, const double *p,
                            const double *qout, double ramp, int Np, double *out);

}  // namespace fpc

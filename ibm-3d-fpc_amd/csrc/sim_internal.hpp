// sim_internal.hpp -- the time-step driver's state, shared by sim.hip and
// adjoint.hip (libfpc).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "fpc_dense.h"
#include "fpc_setup.hpp"
#include "fpc_sim.h"
#include "fpc_struct.h"

namespace fpc {
class NonfiniteCheck;
}

struct fpc_sim_s {
  fpc_params par{};
  hipStream_t stream = nullptr;  // every call on this handle runs on its own stream
  fpc::host::SetupData S;
  // device operator data
  unsigned char *au = nullptr, *av = nullptr, *aw = nullptr, *fp = nullptr, *aqu = nullptr, *aqv = nullptr, *aqw = nullptr;
  double *scu = nullptr, *scv = nullptr, *scw = nullptr, *bcu = nullptr, *dcoef = nullptr, *divbc = nullptr;
  double *uin = nullptr, *uin_step = nullptr;
  // inlet amplitude field theta (ny*nz, 1 = the reference inlet): uin = uin_base * theta and
  // div_bc = div_bc_base * theta (the Dirichlet inlet of BC cleanup and of the divergence)
  std::vector<double> uin_base, divbc_base, theta;
  long long *fu_i = nullptr, *fv_i = nullptr, *px_i = nullptr, *py_i = nullptr;
  double *fu_w = nullptr, *fv_w = nullptr, *px_w = nullptr, *py_w = nullptr;
  fpc_flow_ops ops{};
  // fields
  double *u = nullptr, *v = nullptr, *w = nullptr, *p = nullptr, *ut = nullptr, *vt = nullptr, *wt = nullptr;
  double *qout = nullptr, *diag = nullptr, *hdiag = nullptr;
  // Struct objects
  fpc_StructGrid gu = nullptr, gv = nullptr, gw = nullptr, gp = nullptr;
  fpc_StructStencil st = nullptr;
  fpc_StructMatrix Au = nullptr, Av = nullptr, Aw = nullptr, Ap = nullptr;
  fpc_StructVector bu = nullptr, xu = nullptr, bv = nullptr, xv = nullptr, bw = nullptr, xw = nullptr, bp = nullptr,
                   xp = nullptr;
  fpc_StructSolver su = nullptr, sv = nullptr, sw = nullptr, sp = nullptr, pu = nullptr, pv = nullptr, pw = nullptr,
                   pp = nullptr;
  double *pbu = nullptr, *pxu = nullptr, *pbv = nullptr, *pxv = nullptr, *pbw = nullptr, *pxw = nullptr, *pbp = nullptr,
         *pxp = nullptr;
  int step = 0;
  double ramp = 0.0;
  int its[4] = {0, 0, 0, 0};
  long long its_total[4] = {0, 0, 0, 0};  // summed over the steps run (u, v, w, p)
  double rel[4] = {0, 0, 0, 0};
  int err[4] = {0, 0, 0, 0};
  double t_build = 0, t_mom = 0, t_p = 0, t_proj = 0, t_diag = 0;
  fpc::NonfiniteCheck *nf_mom = nullptr;  // the three momentum solutions, one check and one sync
};

namespace fpc {
// One phase of a step (0 ramp .. 7 diagnostics) on the calling thread's stream.
int sim_run_phase(fpc_sim M, int ph);
}  // namespace fpc

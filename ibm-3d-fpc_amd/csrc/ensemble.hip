// ensemble.hip -- many independent channels at once (libfpc).
//
// No counterpart in the reference (one channel per run, main.cu).  An ensemble
// holds n simulation handles with their own parameters (uneven grids, cylinder
// positions, Reynolds numbers).  Each handle runs on a HIP stream of its own, and
// a pool of persistent host worker threads advances them concurrently: a worker
// takes the next instance (largest first) from a shared counter and runs its
// steps.  The instances' small, latency-bound kernels then overlap on the GPU
// instead of running back to back; every instance computes exactly what it
// computes alone (same kernels, same order of arithmetic).
//
// With nthreads = 0 the ensemble runs batched instead: one stream, and every
// stage of a step is a few launches over all instances.  The momentum solves
// of all instances (3n systems) and their pressure solves (n systems) each run
// as one batch group (batched.hip): one launch per Arnoldi kernel and per
// V-cycle level over every system; the dense stages are enqueued per instance
// back to back on the same stream, with no host synchronisation between them.
// Every instance still computes exactly what it computes alone.
//
// The packed offset table (fpc_ensemble_offsets) lays a field of every instance
// end to end, for callers that post-process the ensemble as one array.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "batched.hpp"
#include "dense_device.hpp"
#include "fpc_internal.hpp"
#include "fpc_setup.hpp"
#include "fpc_sim.h"
#include "sim_internal.hpp"

namespace fpc {
struct BStage;
int stage_create(const std::vector<fpc_sim> &sims, BStage **out);
void stage_destroy(BStage *S);
int stage_pre(BStage *S, const std::vector<fpc_sim> &sims);
int stage_mid(BStage *S);
int stage_post(BStage *S);
}  // namespace fpc

struct fpc_ensemble_s {
  std::vector<fpc_params> par;
  std::vector<fpc_sim> sims;
  std::vector<int> order;  // instances by decreasing cell count
  // worker pool
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv_job, cv_done;
  unsigned long long generation = 0;
  int job = 0;  // 1 create, 2 step
  int nsteps = 0;
  bool stop = false;
  int busy = 0;
  std::atomic<int> next{0};
  std::atomic<int> first_error{0};
  // batched mode
  bool batched = false;
  hipStream_t stream = nullptr;
  fpc::BatchGroup *mom = nullptr, *pres = nullptr;
  fpc::NonfiniteCheck *nf_mom = nullptr, *nf_p = nullptr;
  std::vector<int> its;
  std::vector<double> rel;
  std::vector<int> err;
  // batched dense stages: per-instance descriptors and packed block tables
  fpc::BStage *stage = nullptr;
};

namespace {

void run_job(fpc_ensemble_s *E) {
  const int n = (int)E->order.size();
  for (;;) {
    const int q = E->next.fetch_add(1);
    if (q >= n) break;
    const int i = E->order[(size_t)q];
    int rc = 0;
    if (E->job == 1) {
      rc = fpc_sim_create(&E->par[(size_t)i], &E->sims[(size_t)i]);
    } else if (E->sims[(size_t)i]) {
      rc = fpc_sim_step(E->sims[(size_t)i], E->nsteps);
      rc &= ~FPC_ERROR_CONV;
    }
    if (rc) {
      int zero = 0;
      E->first_error.compare_exchange_strong(zero, rc);
    }
  }
}

void worker_main(fpc_ensemble_s *E) {
  unsigned long long seen = 0;
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(E->mu);
      E->cv_job.wait(lk, [&] { return E->stop || E->generation != seen; });
      if (E->stop) break;
      seen = E->generation;
    }
    run_job(E);
    std::lock_guard<std::mutex> lk(E->mu);
    if (--E->busy == 0) E->cv_done.notify_all();
  }
  fpc::release_thread_state();
}

// Runs job on every instance with the pool; returns the first error code.
int dispatch(fpc_ensemble_s *E, int job, int nsteps) {
  std::unique_lock<std::mutex> lk(E->mu);
  E->job = job;
  E->nsteps = nsteps;
  E->next.store(0);
  E->first_error.store(0);
  E->busy = (int)E->workers.size();
  E->generation += 1;
  E->cv_job.notify_all();
  E->cv_done.wait(lk, [&] { return E->busy == 0; });
  return E->first_error.load();
}

}  // namespace

namespace fpc {

// One instance as the batched dense stages see it.
struct BInst {
  fpc_flow_ops o;
  double *u, *v, *w, *p, *ut, *vt, *wt, *qout, *diag, *uin_step;
  const double *uin;
  double *rb[3], *xb[3];
  double *pbp, *pxp;
  int zero, Np;      // Np: fluid cells (the divL2 mean)
  long long ncell;   // nx * ny * nz
};


struct BTable {
  int *item = nullptr, *first = nullptr;
  int nblocks = 0;
};

struct BStage {
  int n = 0;
  BInst *inst = nullptr;
  BTable inlet, faces, rows, scat, cells;
  double *ramp = nullptr, *hramp = nullptr;     // per instance, device / pinned host
  double *part = nullptr;                       // n * 6 * kDiagBlocks
  int diag_nb = 1;                              // most diagnostics blocks of an instance
  double *ddiag = nullptr, *hdiag = nullptr;    // n * 8, device / pinned host
};

namespace {

#define BS_POINT(count)                                              \
  const int q = item[blockIdx.x];                                    \
  const long long t = (long long)(blockIdx.x - first[q]) * 256 + threadIdx.x; \
  if (t >= (count)) return

__device__ inline long long faces_of(const fpc_flow_ops &o) {
  const long long Nu = (long long)(o.nx + 1) * o.ny * o.nz, Nv = (long long)o.nx * (o.ny + 1) * o.nz,
                  Nw = (long long)o.nx * o.ny * (o.nz + 1);
  return Nu > Nv ? (Nu > Nw ? Nu : Nw) : (Nv > Nw ? Nv : Nw);
}

__global__ void kb_inlet(const BInst *I, const int *item, const int *first, const double *ramp) {
  BS_POINT((long long)I[q].o.ny * I[q].o.nz);
  I[q].uin_step[t] = ramp[q] * I[q].uin[t];
}

// which 0: BC cleanup of u, v, w; 1: of u~, v~, w~
__global__ void kb_bc(const BInst *I, const int *item, const int *first, int which) {
  BS_POINT(faces_of(I[q].o));
  const BInst &B = I[q];
  const fpc_flow_ops &o = B.o;
  if (which == 0)
    bc_point(t, B.u, B.v, B.w, o.active_u, o.active_v, o.active_w, B.uin_step, o.nx, o.ny, o.nz);
  else
    bc_point(t, B.ut, B.vt, B.wt, o.active_u, o.active_v, o.active_w, B.uin_step, o.nx, o.ny, o.nz);
}

// items 3 q + c: momentum box rows of component c of instance q
__global__ void kb_explicit(const BInst *I, const int *item, const int *first) {
  const int e = item[blockIdx.x];
  const long long t = (long long)(blockIdx.x - first[e]) * 256 + threadIdx.x;
  const BInst &B = I[e / 3];
  const int c = e % 3;
  const int bx = B.o.nx - (c == 0), by = B.o.ny - (c == 1), bz = B.o.nz - (c == 2);
  if (t >= (long long)bx * by * bz) return;
  const int i = (int)(t % bx), j = (int)((t / bx) % by), k = (int)(t / ((long long)bx * by));
  explicit_row(c, B.o, B.u, B.v, B.w, B.p, B.rb[c], B.xb[c], B.zero, i, j, k);
}

// items 3 q + c: full face array of component c of instance q
__global__ void kb_scatter(const BInst *I, const int *item, const int *first) {
  const int e = item[blockIdx.x];
  const long long t = (long long)(blockIdx.x - first[e]) * 256 + threadIdx.x;
  const BInst &B = I[e / 3];
  const int c = e % 3;
  double *out = c == 0 ? B.ut : (c == 1 ? B.vt : B.wt);
  const double *base = c == 0 ? B.u : (c == 1 ? B.v : B.w);
  scatter_point(c, t, out, B.xb[c], base, B.o.nx, B.o.ny, B.o.nz);
}

__global__ void kb_divergence(const BInst *I, const int *item, const int *first, const double *ramp) {
  BS_POINT(I[q].ncell);
  const BInst &B = I[q];
  const int nx = B.o.nx, ny = B.o.ny;
  divergence_cell(B.o, B.ut, B.vt, B.wt, ramp[q], B.pbp, B.pxp, (int)(t % nx), (int)((t / nx) % ny),
                  (int)(t / ((long long)nx * ny)));
}

__global__ void kb_project(const BInst *I, const int *item, const int *first) {
  BS_POINT(faces_of(I[q].o));
  const BInst &B = I[q];
  project_point(B.o, t, B.pxp, B.ut, B.vt, B.wt, B.u, B.v, B.w, B.qout);
}

__global__ void kb_add_phi(const BInst *I, const int *item, const int *first) {
  BS_POINT(I[q].ncell);
  const BInst &B = I[q];
  if (B.o.fluid_p[t]) B.p[t] += B.pxp[t];
}

__global__ void __launch_bounds__(256) kb_diag_partials(const BInst *I, const double *ramp, double *part) {
  const BInst &B = I[blockIdx.y];
  const int nb = diag_blocks(B.o.nx, B.o.ny, B.o.nz);  // the partition of stage_diagnostics_async
  if ((int)blockIdx.x >= nb) return;
  diag_partials_body(B.o, B.u, B.v, B.w, B.p, B.qout, ramp[blockIdx.y], part + (long long)blockIdx.y * 6 * kDiagBlocks,
                     blockIdx.x, nb);
}

__global__ void __launch_bounds__(1024) kb_diag_final(const BInst *I, const double *part, double *out) {
  const BInst &B = I[blockIdx.x];
  double *o = out + (long long)blockIdx.x * 8;
  diag_final_body(part + (long long)blockIdx.x * 6 * kDiagBlocks, diag_blocks(B.o.nx, B.o.ny, B.o.nz), B.Np, o);
  __syncthreads();
  if (threadIdx.x < 6) B.diag[threadIdx.x] = o[threadIdx.x];
}

int btable(const std::vector<long long> &count, BTable &t) {
  std::vector<int> item, first(count.size(), 0);
  for (size_t e = 0; e < count.size(); ++e) {
    first[e] = (int)item.size();
    for (long long b = 0; b < (count[e] + 255) / 256; ++b) item.push_back((int)e);
  }
  t.nblocks = (int)item.size();
  FPC_HIP(hipMalloc(&t.item, std::max<size_t>(1, item.size()) * sizeof(int)));
  FPC_HIP(hipMalloc(&t.first, std::max<size_t>(1, first.size()) * sizeof(int)));
  if (!item.empty()) FPC_HIP(hipMemcpy(t.item, item.data(), item.size() * sizeof(int), hipMemcpyHostToDevice));
  FPC_HIP(hipMemcpy(t.first, first.data(), first.size() * sizeof(int), hipMemcpyHostToDevice));
  return 0;
}

void bfree(BTable &t) {
  if (t.item) (void)hipFree(t.item);
  if (t.first) (void)hipFree(t.first);
  t = BTable{};
}

}  // namespace

void stage_destroy(BStage *S) {
  if (!S) return;
  for (BTable *t : {&S->inlet, &S->faces, &S->rows, &S->scat, &S->cells}) bfree(*t);
  if (S->inst) (void)hipFree(S->inst);
  if (S->ramp) (void)hipFree(S->ramp);
  if (S->part) (void)hipFree(S->part);
  if (S->ddiag) (void)hipFree(S->ddiag);
  if (S->hramp) (void)hipHostFree(S->hramp);
  if (S->hdiag) (void)hipHostFree(S->hdiag);
  delete S;
}

int stage_create(const std::vector<fpc_sim> &sims, BStage **out) {
  BStage *S = new BStage();
  const int n = (int)sims.size();
  S->n = n;
  std::vector<BInst> hi((size_t)n);
  std::vector<long long> inlet, faces, rows, scat, cells;
  for (int q = 0; q < n; ++q) {
    fpc_sim M = sims[(size_t)q];
    BInst &B = hi[(size_t)q];
    B.o = M->ops;
    B.u = M->u; B.v = M->v; B.w = M->w; B.p = M->p; B.ut = M->ut; B.vt = M->vt; B.wt = M->wt;
    B.qout = M->qout; B.diag = M->diag; B.uin_step = M->uin_step; B.uin = M->uin;
    B.rb[0] = M->pbu; B.rb[1] = M->pbv; B.rb[2] = M->pbw;
    B.xb[0] = M->pxu; B.xb[1] = M->pxv; B.xb[2] = M->pxw;
    B.pbp = M->pbp; B.pxp = M->pxp;
    B.zero = M->par.use_zero_initial_guess != 0;
    B.Np = M->S.Np_;
    const long long nx = M->par.nx, ny = M->par.ny, nz = M->par.nz;
    B.ncell = nx * ny * nz;
    const long long Nu = (nx + 1) * ny * nz, Nv = nx * (ny + 1) * nz, Nw = nx * ny * (nz + 1);
    inlet.push_back(ny * nz);
    faces.push_back(std::max(Nu, std::max(Nv, Nw)));
    rows.push_back((nx - 1) * ny * nz);
    rows.push_back(nx * (ny - 1) * nz);
    rows.push_back(nx * ny * (nz - 1));
    scat.push_back(Nu);
    scat.push_back(Nv);
    scat.push_back(Nw);
    cells.push_back(nx * ny * nz);
    S->diag_nb = std::max(S->diag_nb, diag_blocks((int)nx, (int)ny, (int)nz));
  }
  int rc = btable(inlet, S->inlet);
  if (!rc) rc = btable(faces, S->faces);
  if (!rc) rc = btable(rows, S->rows);
  if (!rc) rc = btable(scat, S->scat);
  if (!rc) rc = btable(cells, S->cells);
  if (!rc && (hipMalloc(&S->inst, (size_t)n * sizeof(BInst)) != hipSuccess ||
              hipMalloc(&S->ramp, (size_t)n * sizeof(double)) != hipSuccess ||
              hipMalloc(&S->part, (size_t)n * 6 * kDiagBlocks * sizeof(double)) != hipSuccess ||
              hipMalloc(&S->ddiag, (size_t)n * 8 * sizeof(double)) != hipSuccess ||
              hipHostMalloc(&S->hramp, (size_t)n * sizeof(double)) != hipSuccess ||
              hipHostMalloc(&S->hdiag, (size_t)n * 8 * sizeof(double)) != hipSuccess)) {
    set_error(FPC_ERROR_MEMORY, "batched stages: allocation failed");
    rc = FPC_ERROR_MEMORY;
  }
  if (!rc && hipMemcpy(S->inst, hi.data(), (size_t)n * sizeof(BInst), hipMemcpyHostToDevice) != hipSuccess)
    rc = FPC_ERROR_GENERIC;
  if (!rc && hipMemset(S->ddiag, 0, (size_t)n * 8 * sizeof(double)) != hipSuccess) rc = FPC_ERROR_GENERIC;
  if (rc) {
    stage_destroy(S);
    return rc;
  }
  *out = S;
  return 0;
}

// Phase 0 (ramp, inlet) and phase 1 (BC cleanup, momentum rhs and x0) of every instance.
int stage_pre(BStage *S, const std::vector<fpc_sim> &sims) {
  for (int q = 0; q < S->n; ++q) {
    fpc_sim M = sims[(size_t)q];
    M->step += 1;  // main.cu:250-251
    M->ramp = std::min(1.0, (double)M->step / (double)M->par.inlet_ramp_steps);
    S->hramp[q] = M->ramp;
  }
  FPC_HIP(hipMemcpyAsync(S->ramp, S->hramp, (size_t)S->n * sizeof(double), hipMemcpyHostToDevice, stream()));
  kb_inlet<<<S->inlet.nblocks, 256, 0, stream()>>>(S->inst, S->inlet.item, S->inlet.first, S->ramp);
  FPC_LAUNCH_CHECK();
  kb_bc<<<S->faces.nblocks, 256, 0, stream()>>>(S->inst, S->faces.item, S->faces.first, 0);
  FPC_LAUNCH_CHECK();
  kb_explicit<<<S->rows.nblocks, 256, 0, stream()>>>(S->inst, S->rows.item, S->rows.first);
  FPC_LAUNCH_CHECK();
  return 0;
}

// Phases 3 and 4: scatter, BC cleanup of the tilde fields, divergence and pressure rhs.
int stage_mid(BStage *S) {
  kb_scatter<<<S->scat.nblocks, 256, 0, stream()>>>(S->inst, S->scat.item, S->scat.first);
  FPC_LAUNCH_CHECK();
  kb_bc<<<S->faces.nblocks, 256, 0, stream()>>>(S->inst, S->faces.item, S->faces.first, 1);
  FPC_LAUNCH_CHECK();
  kb_divergence<<<S->cells.nblocks, 256, 0, stream()>>>(S->inst, S->cells.item, S->cells.first, S->ramp);
  FPC_LAUNCH_CHECK();
  return 0;
}

// Phases 6 and 7: projection, BC cleanup, p += phi, diagnostics (copied to the host).
int stage_post(BStage *S) {
  kb_project<<<S->faces.nblocks, 256, 0, stream()>>>(S->inst, S->faces.item, S->faces.first);
  FPC_LAUNCH_CHECK();
  kb_bc<<<S->faces.nblocks, 256, 0, stream()>>>(S->inst, S->faces.item, S->faces.first, 0);
  FPC_LAUNCH_CHECK();
  kb_add_phi<<<S->cells.nblocks, 256, 0, stream()>>>(S->inst, S->cells.item, S->cells.first);
  FPC_LAUNCH_CHECK();
  kb_diag_partials<<<dim3(S->diag_nb, S->n), 256, 0, stream()>>>(S->inst, S->ramp, S->part);
  FPC_LAUNCH_CHECK();
  kb_diag_final<<<S->n, 1024, 0, stream()>>>(S->inst, S->part, S->ddiag);
  FPC_LAUNCH_CHECK();
  FPC_HIP(hipMemcpyAsync(S->hdiag, S->ddiag, (size_t)S->n * 8 * sizeof(double), hipMemcpyDeviceToHost, stream()));
  return 0;
}

}  // namespace fpc

namespace {

void stop_workers(fpc_ensemble_s *E) {
  {
    std::lock_guard<std::mutex> lk(E->mu);
    E->stop = true;
    E->cv_job.notify_all();
  }
  for (auto &t : E->workers) t.join();
  E->workers.clear();
}

// Batch groups over the instances' solvers: u, v, w of every instance, and p.
int build_batched(fpc_ensemble_s *E) {
  using namespace fpc;
  if (hipStreamCreate(&E->stream) != hipSuccess) {
    set_error(FPC_ERROR_GENERIC, "fpc_ensemble_create: hipStreamCreate failed");
    return FPC_ERROR_GENERIC;
  }
  StreamScope scope(E->stream);
  std::vector<fpc_StructSolver> sm, sp;
  std::vector<const double *> bm, bp, xm_c, xp_c;
  std::vector<double *> xm, xp;
  std::vector<long long> lm, lp;
  for (fpc_sim M : E->sims) {
    const long long nx = M->par.nx, ny = M->par.ny, nz = M->par.nz;
    const fpc_StructSolver s3[3] = {M->su, M->sv, M->sw};
    const double *b3[3] = {M->pbu, M->pbv, M->pbw};
    double *x3[3] = {M->pxu, M->pxv, M->pxw};
    const long long n3[3] = {(nx - 1) * ny * nz, nx * (ny - 1) * nz, nx * ny * (nz - 1)};
    for (int c = 0; c < 3; ++c) {
      sm.push_back(s3[c]);
      bm.push_back(b3[c]);
      xm.push_back(x3[c]);
      xm_c.push_back(x3[c]);
      lm.push_back(n3[c]);
    }
    sp.push_back(M->sp);
    bp.push_back(M->pbp);
    xp.push_back(M->pxp);
    xp_c.push_back(M->pxp);
    lp.push_back(nx * ny * nz);
  }
  int rc = batch_group_create(sm, bm, xm, &E->mom);
  if (!rc) rc = batch_group_create(sp, bp, xp, &E->pres);
  if (!rc) rc = nonfinite_create(xm_c, lm, &E->nf_mom);
  if (!rc) rc = nonfinite_create(xp_c, lp, &E->nf_p);
  if (!rc) rc = stage_create(E->sims, &E->stage);
  E->its.assign(sm.size(), 0);
  E->rel.assign(sm.size(), 0.0);
  E->err.assign(sm.size(), 0);
  return rc;
}

// One step of every instance: main.cu:249-376 per instance, in batched launches.
int batched_step(fpc_ensemble_s *E) {
  using namespace fpc;
  StreamScope scope(E->stream);
  const int n = (int)E->sims.size();
  int rc = stage_pre(E->stage, E->sims);
  if (rc) return rc;
  // the x0 just written: zero where the instance asks for it (main.cu:272), so the
  // solves copy b as their initial residual
  for (fpc_sim M : E->sims)
    if (M->par.use_zero_initial_guess)
      for (fpc_StructSolver s : {M->su, M->sv, M->sw}) gmres_next_guess_is_zero(s);
  if ((rc = batch_group_solve(E->mom, E->its.data(), E->rel.data(), E->err.data()))) return rc;
  for (int i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) {
      E->sims[(size_t)i]->its[c] = E->its[(size_t)(3 * i + c)];
      E->sims[(size_t)i]->rel[c] = E->rel[(size_t)(3 * i + c)];
      E->sims[(size_t)i]->err[c] = E->err[(size_t)(3 * i + c)];
    }
  if ((rc = nonfinite_enqueue(E->nf_mom))) return rc;
  FPC_HIP(hipStreamSynchronize(stream()));
  for (int i = 0; i < n; ++i) {
    const int *f = nonfinite_flags(E->nf_mom) + 3 * i;
    if (f[0] | f[1] | f[2]) {
      set_error(FPC_ERROR_GENERIC, "non-finite momentum solve output at step %d of instance %d (reference exit 911)",
                E->sims[(size_t)i]->step, i);
      return 911;
    }
  }
  if ((rc = stage_mid(E->stage))) return rc;
  for (fpc_sim M : E->sims) gmres_next_guess_is_zero(M->sp);  // pressure x0 = 0 (main.cu:320)
  if ((rc = batch_group_solve(E->pres, E->its.data(), E->rel.data(), E->err.data()))) return rc;
  for (int i = 0; i < n; ++i) {
    E->sims[(size_t)i]->its[3] = E->its[(size_t)i];
    E->sims[(size_t)i]->rel[3] = E->rel[(size_t)i];
    E->sims[(size_t)i]->err[3] = E->err[(size_t)i];
  }
  if ((rc = nonfinite_enqueue(E->nf_p))) return rc;
  if ((rc = stage_post(E->stage))) return rc;
  FPC_HIP(hipStreamSynchronize(stream()));
  for (int i = 0; i < n; ++i) {
    std::memcpy(E->sims[(size_t)i]->hdiag, E->stage->hdiag + 8 * i, 8 * sizeof(double));
    for (int c = 0; c < 4; ++c) E->sims[(size_t)i]->its_total[c] += E->sims[(size_t)i]->its[c];
  }
  for (int i = 0; i < n; ++i)
    if (nonfinite_flags(E->nf_p)[i]) {
      set_error(FPC_ERROR_GENERIC, "non-finite pressure solve output at step %d of instance %d (reference exit 912)",
                E->sims[(size_t)i]->step, i);
      return 912;
    }
  return 0;
}

long long field_len(const fpc_params &p, const char *field) {
  const long long nx = p.nx, ny = p.ny, nz = p.nz;
  if (!std::strcmp(field, "u")) return (nx + 1) * ny * nz;
  if (!std::strcmp(field, "v")) return nx * (ny + 1) * nz;
  if (!std::strcmp(field, "w")) return nx * ny * (nz + 1);
  if (!std::strcmp(field, "p")) return nx * ny * nz;
  return -1;
}

}  // namespace

extern "C" {

int fpc_ensemble_create(const fpc_params *pars, int n, int nthreads, fpc_ensemble *out) {
  FPC_CHECK_ARG(pars && out && n > 0, "bad arguments");
  FPC_CHECK_ARG(nthreads >= 0 && nthreads <= 64, "nthreads must be 0 (batched) or 1..64");
  FPC_CHECK_ARG(nthreads > 0 || n <= 256, "a batched ensemble holds at most 256 instances");
  fpc::ensure_init();
  fpc_ensemble E = new fpc_ensemble_s();
  E->batched = (nthreads == 0);
  if (E->batched) nthreads = std::min(8, n);  // host setup of the instances runs in parallel
  E->par.assign(pars, pars + n);
  E->sims.assign((size_t)n, nullptr);
  E->order.resize((size_t)n);
  std::iota(E->order.begin(), E->order.end(), 0);
  std::stable_sort(E->order.begin(), E->order.end(), [&](int a, int b) {
    const fpc_params &A = E->par[(size_t)a], &B = E->par[(size_t)b];
    return (long long)A.nx * A.ny * A.nz > (long long)B.nx * B.ny * B.nz;
  });
  for (int t = 0; t < nthreads; ++t) E->workers.emplace_back(worker_main, E);
  int rc = dispatch(E, 1, 0);
  if (rc) {
    fpc_ensemble_destroy(E);
    fpc::set_error(rc, "fpc_ensemble_create: an instance failed to set up");
    return rc;
  }
  if (E->batched) {
    stop_workers(E);
    rc = build_batched(E);
    if (rc) {
      fpc_ensemble_destroy(E);
      return rc;
    }
  }
  *out = E;
  return 0;
}

int fpc_ensemble_step(fpc_ensemble E, int nsteps) {
  FPC_CHECK_ARG(E && nsteps >= 0, "bad arguments");
  FPC_CHECK_ARG(E->batched || E->workers.size() <= 1 || !fpc::prof_enabled(),
                "kernel timing (fpc_ProfileEnable) is for one thread: disable it to step a multi-threaded ensemble");
  if (!E->batched) return dispatch(E, 2, nsteps);
  for (int s = 0; s < nsteps; ++s) {
    int rc = batched_step(E);
    if (rc) return rc;
  }
  return 0;
}

int fpc_ensemble_size(fpc_ensemble E) { return E ? (int)E->sims.size() : 0; }

int fpc_ensemble_tiles(fpc_ensemble E, int *kz, long long *tile_min) {
  FPC_CHECK_ARG(E && kz && tile_min, "null argument");
  FPC_CHECK_ARG(E->batched && E->pres, "not a batched ensemble");
  fpc::batch_group_tiles(E->pres, kz, tile_min);
  return 0;
}

fpc_sim fpc_ensemble_member(fpc_ensemble E, int i) {
  if (!E || i < 0 || i >= (int)E->sims.size()) return nullptr;
  return E->sims[(size_t)i];
}

int fpc_ensemble_offsets(fpc_ensemble E, const char *field, long long *offsets) {
  FPC_CHECK_ARG(E && field && offsets, "null argument");
  offsets[0] = 0;
  for (size_t i = 0; i < E->par.size(); ++i) {
    const long long len = field_len(E->par[i], field);
    FPC_CHECK_ARG(len >= 0, "field must be u, v, w or p");
    offsets[i + 1] = offsets[i] + len;
  }
  return 0;
}

int fpc_ensemble_pack(fpc_ensemble E, const char *field, double *dst, long long nbytes) {
  FPC_CHECK_ARG(E && field && dst, "null argument");
  std::vector<long long> off(E->par.size() + 1);
  int rc = fpc_ensemble_offsets(E, field, off.data());
  if (rc) return rc;
  FPC_CHECK_ARG(nbytes >= off.back() * (long long)sizeof(double), "destination too small");
  for (size_t i = 0; i < E->sims.size(); ++i) {
    rc = fpc_sim_copy(E->sims[i], field, dst + off[i], (off[i + 1] - off[i]) * (long long)sizeof(double));
    if (rc) return rc;
  }
  return 0;
}

int fpc_ensemble_destroy(fpc_ensemble E) {
  if (!E) return 0;
  stop_workers(E);
  if (E->mom) fpc::batch_group_destroy(E->mom);
  if (E->pres) fpc::batch_group_destroy(E->pres);
  if (E->nf_mom) fpc::nonfinite_destroy(E->nf_mom);
  if (E->nf_p) fpc::nonfinite_destroy(E->nf_p);
  if (E->stage) fpc::stage_destroy(E->stage);
  if (E->stream) (void)hipStreamDestroy(E->stream);
  for (fpc_sim s : E->sims)
    if (s) fpc_sim_destroy(s);
  delete E;
  return 0;
}

}  // extern "C"

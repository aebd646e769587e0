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
#include "fpc_internal.hpp"
#include "fpc_setup.hpp"
#include "fpc_sim.h"
#include "sim_internal.hpp"

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
  int rc = 0;
  for (fpc_sim M : E->sims)
    if ((rc = sim_run_phase(M, 0))) return rc;
  for (fpc_sim M : E->sims)
    if ((rc = stage_explicit_async(&M->ops, M->u, M->v, M->w, M->p, M->uin_step, M->pbu, M->pxu, M->pbv, M->pxv,
                                   M->pbw, M->pxw, M->par.use_zero_initial_guess != 0)))
      return rc;
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
  for (fpc_sim M : E->sims)
    if ((rc = sim_run_phase(M, 3)) || (rc = sim_run_phase(M, 4))) return rc;
  if ((rc = batch_group_solve(E->pres, E->its.data(), E->rel.data(), E->err.data()))) return rc;
  for (int i = 0; i < n; ++i) {
    E->sims[(size_t)i]->its[3] = E->its[(size_t)i];
    E->sims[(size_t)i]->rel[3] = E->rel[(size_t)i];
    E->sims[(size_t)i]->err[3] = E->err[(size_t)i];
  }
  if ((rc = nonfinite_enqueue(E->nf_p))) return rc;
  for (fpc_sim M : E->sims) {
    if ((rc = stage_projection_async(&M->ops, M->pxp, M->ut, M->vt, M->wt, M->u, M->v, M->w, M->p, M->uin_step,
                                     M->qout)))
      return rc;
    if ((rc = stage_diagnostics_async(&M->ops, M->u, M->v, M->w, M->p, M->qout, M->ramp, M->S.Np_, M->diag)))
      return rc;
    FPC_HIP(hipMemcpyAsync(M->hdiag, M->diag, 8 * sizeof(double), hipMemcpyDeviceToHost, stream()));
  }
  FPC_HIP(hipStreamSynchronize(stream()));
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
  if (E->stream) (void)hipStreamDestroy(E->stream);
  for (fpc_sim s : E->sims)
    if (s) fpc_sim_destroy(s);
  delete E;
  return 0;
}

}  // extern "C"

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

#include "fpc_internal.hpp"
#include "fpc_sim.h"

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
  FPC_CHECK_ARG(nthreads >= 1 && nthreads <= 64, "nthreads must be 1..64");
  fpc::ensure_init();
  fpc_ensemble E = new fpc_ensemble_s();
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
  *out = E;
  return 0;
}

int fpc_ensemble_step(fpc_ensemble E, int nsteps) {
  FPC_CHECK_ARG(E && nsteps >= 0, "bad arguments");
  return dispatch(E, 2, nsteps);
}

int fpc_ensemble_size(fpc_ensemble E) { return E ? (int)E->sims.size() : 0; }

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
  {
    std::lock_guard<std::mutex> lk(E->mu);
    E->stop = true;
    E->cv_job.notify_all();
  }
  for (auto &t : E->workers) t.join();
  for (fpc_sim s : E->sims)
    if (s) fpc_sim_destroy(s);
  delete E;
  return 0;
}

}  // extern "C"

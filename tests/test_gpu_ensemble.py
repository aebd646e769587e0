"""GPU: an ensemble of uneven channels, stepped concurrently (one HIP stream per
instance, several host worker threads) or batched (one stream, one launch per
stage over every instance), computes exactly what each instance computes
alone, and its packed offset table lays the fields end to end."""
import numpy as np
import pytest

import fpc_amd as F
from helpers import bits_equal

pytestmark = pytest.mark.gpu


def _params():
    return F.ensemble_workload(n=6, seed=3, nx=(24, 48), nyz=(8, 14), inlet_ramp_steps=4)


@pytest.mark.parametrize("nthreads", [3, 0], ids=["concurrent", "batched"])
def test_ensemble_equals_instances_run_alone(torch_cuda, nthreads):
    pars = _params()
    ens = F.Ensemble(pars, nthreads=nthreads)
    ens.step(3)
    for i, p in enumerate(pars):
        alone = F.Sim(p)
        alone.step(3)
        for f in ("u", "v", "w", "p"):
            assert bits_equal(ens.members[i].get(f), alone.get(f)), (i, f)
        for k in ("its_u", "its_v", "its_w", "its_p", "relres_p", "cd", "cl", "div_l2"):
            assert ens.members[i].scalar(k) == alone.scalar(k), (i, k)
        alone.close()
    ens.close()


def test_batched_ensemble_restarts_and_tight_tolerance(torch_cuda):
    """Short restart length (kdim 4) and rtol 1e-12: every system restarts several
    times at its own pace, and the batch still reproduces each instance bitwise."""
    pars = F.ensemble_workload(n=5, seed=8, nx=(24, 40), nyz=(8, 12), inlet_ramp_steps=3, vel_kdim=4, p_kdim=4,
                               vel_tol=1e-12, p_tol=1e-12, vel_maxit=400, p_maxit=400)
    ens = F.Ensemble(pars, nthreads=0)
    ens.step(2)
    for i, p in enumerate(pars):
        alone = F.Sim(p)
        alone.step(2)
        assert alone.scalar("its_p") > 4
        for f in ("u", "v", "w", "p"):
            assert bits_equal(ens.members[i].get(f), alone.get(f)), (i, f)
        for k in ("its_u", "its_v", "its_w", "its_p"):
            assert ens.members[i].scalar(k) == alone.scalar(k), (i, k)
        alone.close()
    ens.close()


def test_ensemble_packed_offsets(torch_cuda):
    pars = _params()
    ens = F.Ensemble(pars, nthreads=2)
    ens.step(1)
    off = ens.offsets("p")
    assert off[0] == 0 and len(off) == len(pars) + 1
    assert list(np.diff(off)) == [p.nx * p.ny * p.nz for p in pars]
    packed = ens.pack("p")
    for i in range(len(pars)):
        assert bits_equal(packed[off[i]:off[i + 1]], ens.members[i].get("p"))
    ens.close()


def test_batched_ensemble_tiled_instances(torch_cuda):
    """Instances large enough that the batch runs their matvecs and V-cycle levels
    on z-marching tiles (>= 16384 points, nx >= 32, ny >= 8), while each instance
    alone runs one-point kernels: the two kernel forms give the same bits."""
    pars = F.ensemble_workload(n=4, seed=12, nx=(64, 96), nyz=(16, 24), inlet_ramp_steps=3)
    ens = F.Ensemble(pars, nthreads=0)
    ens.step(2)
    for i, p in enumerate(pars):
        alone = F.Sim(p)
        alone.step(2)
        for f in ("u", "v", "w", "p"):
            assert bits_equal(ens.members[i].get(f), alone.get(f)), (i, f)
        for k in ("its_u", "its_v", "its_w", "its_p"):
            assert ens.members[i].scalar(k) == alone.scalar(k), (i, k)
        alone.close()
    ens.close()


def test_threaded_ensemble_refuses_kernel_timing(torch_cuda):
    """Kernel timing keeps process-wide records for one calling thread: a concurrent
    ensemble with several workers refuses to step while it is on, and steps again
    once it is off."""
    pars = F.ensemble_workload(n=3, seed=4, nx=(24, 32), nyz=(8, 10), inlet_ramp_steps=2)
    ens = F.Ensemble(pars, nthreads=2)
    F.profile_enable("*")
    try:
        with pytest.raises(F.FpcError):
            ens.step(1)
    finally:
        F.profile_enable("")
    ens.step(1)
    assert all(m.scalar("step") == 1 for m in ens.members)
    ens.close()


def test_batched_ensemble_with_captured_vcycles(torch_cuda):
    """With "mg_graph" on, the batched V-cycle of each Arnoldi step is captured once and
    replayed: every instance still equals its standalone run bit for bit."""
    pars = _params()
    F.set_tuning("mg_graph", 1)
    try:
        ens = F.Ensemble(pars, nthreads=0)
        ens.step(3)
        got = [{f: m.get(f) for f in ("u", "v", "w", "p")} for m in ens.members]
        ens.close()
    finally:
        F.set_tuning("mg_graph", 0)
    for i, p in enumerate(pars):
        alone = F.Sim(p)
        alone.step(3)
        for f in ("u", "v", "w", "p"):
            assert bits_equal(got[i][f], alone.get(f)), (i, f)
        alone.close()

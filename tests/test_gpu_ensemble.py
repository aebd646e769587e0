"""GPU: an ensemble of uneven channels stepped concurrently (one HIP stream per
instance, several host worker threads) computes exactly what each instance
computes alone, and its packed offset table lays the fields end to end."""
import numpy as np
import pytest

import fpc_amd as F
from helpers import bits_equal

pytestmark = pytest.mark.gpu


def _params():
    return F.ensemble_workload(n=6, seed=3, nx=(24, 48), nyz=(8, 14), inlet_ramp_steps=4)


def test_ensemble_equals_instances_run_alone(torch_cuda):
    pars = _params()
    ens = F.Ensemble(pars, nthreads=3)
    ens.step(3)
    for i, p in enumerate(pars):
        alone = F.Sim(p)
        alone.step(3)
        for f in ("u", "v", "w", "p"):
            assert bits_equal(ens.members[i].get(f), alone.get(f)), (i, f)
        assert ens.members[i].scalar("cd") == alone.scalar("cd")
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

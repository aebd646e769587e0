"""GPU runs at the sizes of the mixed-precision and adjoint workloads.

* 1024x160x160 (26M cells, Um 2.25, Re ~100): three steps in FP64 and with the
  FP32 V-cycle and FP32 Jacobi inside FP64 flexible GMRES.  Both reach the FP64
  1e-8 residuals with the same iteration counts, the fields agree to 1e-10
  (relative L2; the FP32 preconditioner changes only the rounding of converged
  solves), and the projected field is discretely divergence-free to the pressure
  solve's tolerance.
* 256x64x64 (Re ~20): the checkpointed adjoint gradient of mean c_D over steps
  21..30 with respect to the 64x64 inlet amplitude field and nu, against central
  differences (h = 1e-6) of the forward solve on two seeded directions, to 1e-5
  relative; solves to rtol 1e-12 so the differences resolve 1e-7.
"""
import numpy as np
import pytest

import fpc_amd as F
from helpers import rel_l2

pytestmark = pytest.mark.gpu

MIXED = dict(nx=1024, ny=160, nz=160, Lx=2.5, Ly=0.41, Lz=0.41, xcirc=0.5, ycirc=0.2, Rcirc=0.05, Um=2.25, nu=1e-3,
             rho=1.0, dt=2.5e-4, vel_kdim=30, p_kdim=30, vel_tol=1e-8, p_tol=1e-8, vel_maxit=250, p_maxit=250,
             init_profile=1, noise_amp=1e-3 * 2.25, seed=0)
ADJOINT = dict(nx=256, ny=64, nz=64, Lx=2.5, Ly=0.41, Lz=0.41, xcirc=0.5, ycirc=0.2, Rcirc=0.05, Um=0.45, nu=1e-3,
               rho=1.0, dt=1e-3, vel_kdim=30, p_kdim=30, vel_tol=1e-12, p_tol=1e-12, vel_maxit=500, p_maxit=500,
               init_profile=1, noise_amp=1e-3 * 0.45, seed=0)


def _run(kw, nsteps):
    s = F.Sim(F.default_params(**kw))
    rows = []
    for _ in range(nsteps):
        s.step(1)
        rows.append(s.scalars())
    fields = {f: s.get(f) for f in "uvwp"}
    s.close()
    return rows, fields


def test_mixed_precision_at_1024x160x160(torch_cuda):
    r64, f64 = _run(MIXED, 3)
    r32, f32 = _run(dict(MIXED, mg_precision=32), 3)
    for a, b in zip(r64, r32):
        for c in "uvwp":
            assert a["its_" + c] == b["its_" + c], c
        assert max(a["relres_" + c] for c in "uvwp") <= 1e-8
        assert max(b["relres_" + c] for c in "uvwp") <= 1e-8
        assert abs(a["cd"] - b["cd"]) <= 1e-10 * abs(a["cd"])
    for f in "uvwp":
        assert rel_l2(f32[f], f64[f]) < 1e-10, f
    # projection: div_l2 relative to max|u| / h at the pressure solve's tolerance
    h = 0.41 / 160
    assert r64[-1]["div_l2"] < 1e-6 * r64[-1]["umax"] / h


def test_adjoint_gradient_at_256x64x64(torch_cuda):
    nsteps, avg_from, every, h = 30, 21, 10, 1e-6
    nin = ADJOINT["ny"] * ADJOINT["nz"]
    rng = np.random.default_rng(0)
    theta = 1.0 + 0.05 * (2.0 * rng.random(nin) - 1.0)
    nu0 = ADJOINT["nu"]

    def J_of(th, nu):
        s = F.Sim(F.default_params(**dict(ADJOINT, nu=nu)))
        s.set("inlet_amplitude", th)
        cds = []
        for n in range(1, nsteps + 1):
            s.step(1)
            if n >= avg_from:
                cds.append(s.scalar("cd"))
        s.close()
        return float(np.mean(cds))

    sim = F.Sim(F.default_params(**ADJOINT))
    sim.set("inlet_amplitude", theta)
    adj = F.Adjoint(sim)
    J, g_th, g_nu = adj.drag_gradient(nsteps, avg_from, every)
    adj.close()
    sim.close()
    assert abs(J - J_of(theta, nu0)) <= 1e-12 * abs(J)
    for _ in range(2):
        dth, dnu = rng.standard_normal(nin), float(rng.standard_normal())
        ad = float(g_th @ dth + g_nu * nu0 * dnu)
        fd = (J_of(theta + h * dth, nu0 * (1 + h * dnu)) - J_of(theta - h * dth, nu0 * (1 - h * dnu))) / (2 * h)
        assert abs(fd - ad) <= 1e-5 * abs(ad), (fd, ad)

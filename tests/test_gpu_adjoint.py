"""GPU: the discrete adjoint of the time step (C-ABI VJP) and the checkpointed
gradient of time-averaged drag with respect to the inlet amplitude field theta
and nu, checked against central finite differences of this build's forward
solve and of the oracle's (the CPU restatement of the reference path).

The objective is the mean of c_D over steps 7..12 of a 32x12x12 channel
(inlet ramp 4 steps), solves converged to rtol 1e-12 so that finite
differences with h = 1e-6 resolve 1e-7.  Bar: 1e-5 relative per direction.
"""
import numpy as np
import pytest

import fpc_amd as F
import oracle as O
from helpers import TINY

pytestmark = pytest.mark.gpu

BASE = dict(TINY, init_profile=1, noise_amp=1e-3, seed=5, inlet_ramp_steps=4, vel_tol=1e-12, p_tol=1e-12)
NSTEPS, AVG_FROM, EVERY = 12, 7, 4
H = 1e-6


def _theta(seed=11):
    rng = np.random.default_rng(seed)
    return 1.0 + 0.05 * (2.0 * rng.random(TINY["ny"] * TINY["nz"]) - 1.0)


def _J(mod, theta, nu):
    s = mod.Sim(mod.default_params(**dict(BASE, nu=nu)))
    s.set("inlet_amplitude", theta)
    cds = []
    for n in range(1, NSTEPS + 1):
        s.step(1)
        if n >= AVG_FROM:
            cds.append(s.scalar("cd"))
    return float(np.mean(cds))


def _directions(n, seed=21):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        yield rng.standard_normal(TINY["ny"] * TINY["nz"]), float(rng.standard_normal())


def test_step_vjp_dot_product(torch_cuda):
    """<a1, (F(U+e d) - F(U-e d)) / 2e> == <VJP(a1), d> for one step F (quadratic in U,
    so the central difference is exact up to rounding)."""
    torch = torch_cuda
    sim = F.Sim(F.default_params(**BASE))
    sim.set("inlet_amplitude", _theta())
    sim.step(2)
    names = ("u", "v", "w", "p")
    U0 = {f: sim.get(f) for f in names}
    rng = np.random.default_rng(3)
    d = {f: rng.standard_normal(U0[f].size) * 1e-2 for f in names}
    a1 = {f: rng.standard_normal(U0[f].size) for f in names}
    eps = 1e-3

    def run(sign):
        for f in names:
            sim.set(f, U0[f] + sign * eps * d[f])
        sim.set_step(2)
        sim.step(1)
        return {f: sim.get(f) for f in names}

    up, um = run(1.0), run(-1.0)
    lhs = sum(np.dot(a1[f], (up[f] - um[f]) / (2 * eps)) for f in names)
    adj = F.Adjoint(sim)
    dev = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    state = [dev(U0[f]) for f in names]
    abar = [dev(a1[f]) for f in names]
    thb = torch.zeros(TINY["ny"] * TINY["nz"], dtype=torch.float64, device="cuda")
    adj.step_vjp(3, state, abar, thb)
    rhs = sum(np.dot(abar[q].cpu().numpy(), d[f]) for q, f in enumerate(names))
    assert abs(lhs - rhs) <= 1e-7 * abs(lhs), (lhs, rhs)


class DragMean:
    """torch.autograd.Function around the C-ABI gradient (test harness only)."""

    @staticmethod
    def make(torch):
        class Fn(torch.autograd.Function):
            @staticmethod
            def forward(ctx, theta, nu):
                sim = F.Sim(F.default_params(**dict(BASE, nu=float(nu))))
                sim.set("inlet_amplitude", theta.detach().cpu().numpy())
                adj = F.Adjoint(sim)
                J, g_th, g_nu = adj.drag_gradient(NSTEPS, AVG_FROM, EVERY)
                ctx.save_for_backward(torch.from_numpy(g_th), torch.tensor(g_nu, dtype=torch.float64))
                return torch.tensor(J, dtype=torch.float64)

            @staticmethod
            def backward(ctx, gJ):
                g_th, g_nu = ctx.saved_tensors
                return gJ * g_th, gJ * g_nu

        return Fn


def test_drag_gradient_matches_finite_differences(torch_cuda):
    torch = torch_cuda
    theta0, nu0 = _theta(), 1e-3
    theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    nu = torch.tensor(nu0, dtype=torch.float64, requires_grad=True)
    J = DragMean.make(torch).apply(theta, nu)
    J.backward()
    g_th, g_nu = theta.grad.numpy(), float(nu.grad)
    # the gradient run's objective is the plain forward run's
    Jv = float(J.detach())
    assert abs(Jv - _J(F, theta0, nu0)) <= 1e-13 * abs(Jv)
    for dth, dnu in _directions(4):
        ad = g_th @ dth + g_nu * nu0 * dnu
        fd = (_J(F, theta0 + H * dth, nu0 * (1 + H * dnu)) - _J(F, theta0 - H * dth, nu0 * (1 - H * dnu))) / (2 * H)
        assert abs(fd - ad) <= 1e-5 * abs(ad), (fd, ad)


def test_drag_gradient_matches_oracle_finite_differences(torch_cuda):
    theta0, nu0 = _theta(), 1e-3
    sim = F.Sim(F.default_params(**BASE))
    sim.set("inlet_amplitude", theta0)
    J, g_th, g_nu = F.Adjoint(sim).drag_gradient(NSTEPS, AVG_FROM, EVERY)
    assert abs(J - _J(O, theta0, nu0)) <= 1e-9 * abs(J)
    for dth, dnu in _directions(2, seed=33):
        ad = g_th @ dth + g_nu * nu0 * dnu
        fd = (_J(O, theta0 + H * dth, nu0 * (1 + H * dnu)) - _J(O, theta0 - H * dth, nu0 * (1 - H * dnu))) / (2 * H)
        assert abs(fd - ad) <= 1e-5 * abs(ad), (fd, ad)

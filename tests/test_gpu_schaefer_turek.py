"""The drag functional against the published steady benchmark.

Schaefer & Turek's 3D-1Z case (Re = 20): channel 2.5 x 0.41 x 0.41, full-height
cylinder D = 0.1 at (0.5, 0.2), inflow 16 Um y z (H - y)(H - z) / H^4 with
Um = 0.45, nu = 1e-3; c_D = 2 F_x / (rho Ubar^2 D H), Ubar = 4 Um / 9.  The
published reference interval is 6.05 <= c_D <= 6.25 (Schaefer & Turek 1996,
"Benchmark computations of laminar flow around a cylinder").

At 512x96x96 (D = 20 cells) and dt = 1e-3 the solver reaches c_D = 6.1267 at
t = 3 and 6.1234 at t = 6: inside the interval.  The lift coefficient of this
discretisation (about 0.17 against the published 0.008-0.010) is not resolved:
it changes sign between grids and is left untested."""
import pytest

import fpc_amd as F

pytestmark = pytest.mark.gpu

ST_1Z = dict(nx=512, ny=96, nz=96, Lx=2.5, Ly=0.41, Lz=0.41, xcirc=0.5, ycirc=0.2, Rcirc=0.05, Um=0.45, nu=1e-3,
             rho=1.0, dt=1e-3, vel_tol=1e-8, p_tol=1e-8, vel_maxit=250, p_maxit=250, init_profile=1,
             inlet_ramp_steps=1)


def test_steady_re20_drag_in_published_interval(torch_cuda):
    s = F.Sim(F.default_params(**ST_1Z))
    cds = []
    for _ in range(12):  # t = 3 in blocks of 250 steps
        s.step(250)
        cds.append(s.scalar("cd"))
        print(f"t = {len(cds) * 0.25:.2f}  c_D = {cds[-1]:.5f}", flush=True)
    s.close()
    assert 6.05 <= cds[-1] <= 6.25, cds
    # close to steady: the last block moved c_D by less than 5e-3
    assert abs(cds[-1] - cds[-2]) < 5e-3, cds
    # and the approach is monotone from above over the last second
    assert all(a >= b for a, b in zip(cds[-5:], cds[-4:])), cds

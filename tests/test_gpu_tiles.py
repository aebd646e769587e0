"""GPU parity of the z-marching stencil kernels against the CPU oracle.

Boxes large enough for the tiled kernels (nx >= 32, ny >= 8, >= 32768 points,
with the size threshold lowered to that floor), with several z-chunks of the
automatic length and of 4 planes, odd and even nz, and couplings scaled per direction so
the first coarsening direction is x, y or z in turn (z exercises the fused
residual + z-restriction).  Operators are random nonsymmetric M-matrices.
Tolerances as for the other Struct tests: 1e-13 for one matvec, 1e-12 for one
V-cycle, 1e-9 for solutions converged to rtol 1e-11.
"""
import numpy as np
import pytest

import fpc_amd as F
import oracle as O
from helpers import bits_equal, box_csr, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, params=[0, 4], ids=["kz-auto", "kz-4"])
def tiled_everywhere(request, torch_cuda):
    """Tiled kernels on every box the tiles accept (>= 32768 points), automatic and
    shortest z-chunks; library defaults restored afterwards."""
    F.set_tuning("tile_min_points", 0)
    F.set_tuning("stencil_kz", request.param)
    yield
    F.set_tuning("tile_min_points", 1 << 19)
    F.set_tuning("stencil_kz", 0)


def _aniso(nx, ny, nz, scale, seed):
    rng = np.random.default_rng(seed)
    v = np.zeros((nz, ny, nx, 7))
    for s in range(1, 7):
        d = (s - 1) // 2
        v[..., s] = -scale[d] * (0.5 + rng.random((nz, ny, nx)))
    v[:, :, 0, 1] = 0
    v[:, :, -1, 2] = 0
    v[:, 0, :, 3] = 0
    v[:, -1, :, 4] = 0
    v[0, :, :, 5] = 0
    v[-1, :, :, 6] = 0
    v[..., 0] = 1.05 * np.abs(v[..., 1:]).sum(axis=-1) + 0.01
    return v.reshape(-1)


CASES = [  # dims, per-direction coupling scale, expected first coarsening direction
    ((96, 40, 37), (1.0, 1.0, 4.0), 2),
    ((80, 24, 36), (1.0, 1.0, 4.0), 2),
    ((96, 40, 37), (1.0, 4.0, 1.0), 1),
    ((96, 40, 21), (4.0, 1.0, 1.0), 0),
]


@pytest.mark.parametrize("dims,scale,cdir", CASES, ids=["z-odd", "z-even", "y", "x"])
def test_tiled_vcycle_matches_oracle(torch_cuda, dims, scale, cdir):
    nx, ny, nz = dims
    vals = _aniso(nx, ny, nz, scale, seed=nz)
    b = np.random.default_rng(1).standard_normal(nx * ny * nz)
    box = F.StructBox(nx, ny, nz, vals)
    lv = O.mg_levels(nx, ny, nz, vals)
    assert lv[0][1] == cdir
    for relax in (0, 1):
        want = O.precond_apply(nx, ny, nz, vals, b, precond="pfmg", relax_type=relax)
        got = F.pfmg_apply(box, b, relax_type=relax)
        assert rel_l2(got, want) < 1e-12, relax


@pytest.mark.parametrize("dims,scale,cdir", CASES[:2], ids=["z-odd", "z-even"])
def test_tiled_matvec_and_gmres_match_oracle(torch_cuda, dims, scale, cdir):
    nx, ny, nz = dims
    vals = _aniso(nx, ny, nz, scale, seed=nz + 1)
    rng = np.random.default_rng(4)
    x = rng.standard_normal(nx * ny * nz)
    box = F.StructBox(nx, ny, nz, vals)
    assert rel_l2(box.matvec(x), O.struct_matvec(nx, ny, nz, vals, x)) < 1e-13
    b = rng.standard_normal(nx * ny * nz)
    A = box_csr(nx, ny, nz, vals)
    for pc in ("jacobi", "none", "pfmg"):
        x_o, its_o, _, rc_o = O.gmres(nx, ny, nz, vals, b, tol=1e-11, maxit=300, kdim=30, precond=pc)
        sol = F.Solver(box, tol=1e-11, maxit=300, kdim=30, precond=pc)
        x_g, its_g, rel_g, rc_g = sol.solve(b)
        assert rc_o == 0 and rc_g == 0, pc
        assert abs(its_g - its_o) <= 1, (pc, its_g, its_o)
        assert rel_l2(x_g, x_o) < 1e-9, pc
        assert np.linalg.norm(b - A @ x_g) / np.linalg.norm(b) < 2e-11, pc


@pytest.mark.parametrize("dims,scale,cdir", CASES, ids=["z-odd", "z-even", "y", "x"])
def test_tiled_equal_one_point_bitwise(torch_cuda, dims, scale, cdir):
    """The z-marching kernels and the one-point-per-thread kernels give the same bits:
    one V-cycle (every level tiled or none), and GMRES with the fused Jacobi matvec
    and with the V-cycle.  Out-of-box coefficients are left non-zero: both forms
    must skip them, not add them times zero."""
    nx, ny, nz = dims
    v = _aniso(nx, ny, nz, scale, seed=nz + 7).reshape(nz, ny, nx, 7)
    rng = np.random.default_rng(9)
    v[:, :, 0, 1] = -0.3 - rng.random((nz, ny))
    v[:, -1, :, 4] = -0.2 - rng.random((nz, nx))
    v[0, :, :, 5] = -0.1 - rng.random((ny, nx))
    vals = v.reshape(-1)
    b = rng.standard_normal(nx * ny * nz)
    box = F.StructBox(nx, ny, nz, vals)
    out = {}
    for mode, tmin in (("tiled", 0), ("one-point", 1 << 40)):
        F.set_tuning("tile_min_points", tmin)
        y = F.pfmg_apply(box, b)
        xp, itp, _, _ = F.Solver(box, tol=1e-10, maxit=200, kdim=30, precond="pfmg").solve(b)
        xj, itj, _, _ = F.Solver(box, tol=1e-10, maxit=400, kdim=30, precond="jacobi").solve(b)
        out[mode] = (y, xp, itp, xj, itj)
    t, o = out["tiled"], out["one-point"]
    assert bits_equal(t[0], o[0])
    assert t[2] == o[2] and bits_equal(t[1], o[1])
    assert t[4] == o[4] and bits_equal(t[3], o[3])

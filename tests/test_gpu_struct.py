"""GPU parity of the Struct stack (C ABI) against the CPU oracle.

SpMV, GMRES+Jacobi, GMRES+PFMG and a single PFMG V-cycle on the embedded
matrices of the reference setup, at sizes the oracle finishes in seconds.
Tolerances: the GPU contracts a*b+c into FMA and reduces dot products in a
different order than the sequential oracle, so agreement is to rounding:
1e-13 relative for one matvec, 1e-12 for one V-cycle, 1e-9 for solutions
converged to rtol 1e-10/1e-12 (the momentum/pressure matrices are well
conditioned), and equal iteration counts within one.
"""
import numpy as np
import pytest

import fpc_amd as F
import oracle as O
from helpers import SMALL, TINY, bits_equal, box_csr, rel_l2

pytestmark = pytest.mark.gpu


def _setup(kw):
    return O.Setup(O.default_params(**kw))


BOXES = {  # name -> (dims from params, AoS array name)
    "u": (lambda p: (p["nx"] - 1, p["ny"], p["nz"]), "Au_vals"),
    "v": (lambda p: (p["nx"], p["ny"] - 1, p["nz"]), "Av_vals"),
    "w": (lambda p: (p["nx"], p["ny"], p["nz"] - 1), "Aw_vals"),
    "p": (lambda p: (p["nx"], p["ny"], p["nz"]), "Ap_vals"),
}


def _box(kw, which):
    S = _setup(kw)
    dims = BOXES[which][0](kw)
    vals = S.get(BOXES[which][1])
    return S, dims, vals


@pytest.mark.parametrize("which", ["u", "v", "w", "p"])
@pytest.mark.parametrize("kw", [TINY, SMALL], ids=["tiny", "small"])
def test_matvec_matches_oracle(torch_cuda, kw, which):
    S, (nx, ny, nz), vals = _box(kw, which)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(nx * ny * nz)
    y0 = rng.standard_normal(nx * ny * nz)
    box = F.StructBox(nx, ny, nz, vals)
    want = O.struct_matvec(nx, ny, nz, vals, x)
    got = box.matvec(x)
    assert rel_l2(got, want) < 1e-13
    # alpha/beta form of HYPRE_StructMatrixMatvec
    got2 = box.matvec(x, alpha=-1.5, beta=0.5, y=y0)
    assert rel_l2(got2, -1.5 * want + 0.5 * y0) < 1e-13
    # and the scipy form of the same box matrix
    assert rel_l2(got, box_csr(nx, ny, nz, vals) @ x) < 1e-13


@pytest.mark.parametrize("which", ["u", "v", "w"])
def test_gmres_jacobi_matches_oracle(torch_cuda, which):
    S, (nx, ny, nz), vals = _box(SMALL, which)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(nx * ny * nz)
    x_o, its_o, rel_o, rc_o = O.gmres(nx, ny, nz, vals, b, tol=1e-12, maxit=200, kdim=30, precond="jacobi")
    box = F.StructBox(nx, ny, nz, vals)
    sol = F.Solver(box, tol=1e-12, maxit=200, kdim=30, precond="jacobi")
    x_g, its_g, rel_g, rc_g = sol.solve(b)
    assert rc_o == 0 and rc_g == 0
    assert abs(its_g - its_o) <= 1
    assert rel_g <= 1e-12
    assert rel_l2(x_g, x_o) < 1e-10
    A = box_csr(nx, ny, nz, vals)
    assert np.linalg.norm(b - A @ x_g) / np.linalg.norm(b) < 1e-11


@pytest.mark.parametrize("which", ["u", "w"])
def test_gmres_fp32_jacobi_reaches_fp64_residual(torch_cuda, which):
    """Jacobi on the stored FP32 1/diag (the mixed-precision momentum preconditioner):
    the preconditioner is still a fixed diagonal, so plain GMRES converges to the same
    FP64 tolerance; solutions agree with the FP64-Jacobi solve and the oracle's to the
    accuracy the tolerance fixes."""
    S, (nx, ny, nz), vals = _box(SMALL, which)
    b = np.random.default_rng(4).standard_normal(nx * ny * nz)
    x_o, its_o, _, _ = O.gmres(nx, ny, nz, vals, b, tol=1e-12, maxit=200, kdim=30, precond="jacobi")
    box = F.StructBox(nx, ny, nz, vals)
    x32, its32, rel32, rc32 = F.Solver(box, tol=1e-12, maxit=200, kdim=30, precond="jacobi", jacobi_bits=32).solve(b)
    assert rc32 == 0 and rel32 <= 1e-12
    A = box_csr(nx, ny, nz, vals)
    assert np.linalg.norm(b - A @ x32) / np.linalg.norm(b) < 1e-11
    assert rel_l2(x32, x_o) < 1e-10
    assert abs(its32 - its_o) <= 2


@pytest.mark.parametrize("relax", [1, 2])
def test_gmres_pfmg_matches_oracle(torch_cuda, relax):
    S, (nx, ny, nz), vals = _box(SMALL, "p")
    fl = S.get("fluid_p").astype(bool)
    rng = np.random.default_rng(5)
    b = rng.standard_normal(nx * ny * nz)
    b[~fl] = 0.0
    x_o, its_o, rel_o, rc_o = O.gmres(nx, ny, nz, vals, b, tol=1e-10, maxit=200, kdim=30, precond="pfmg",
                                      relax_type=relax)
    box = F.StructBox(nx, ny, nz, vals)
    sol = F.Solver(box, tol=1e-10, maxit=200, kdim=30, precond="pfmg", relax_type=relax)
    x_g, its_g, rel_g, rc_g = sol.solve(b)
    assert rc_o == 0 and rc_g == 0, (rc_o, rc_g, its_o, its_g)
    assert abs(its_g - its_o) <= 1, (its_g, its_o)
    assert rel_l2(x_g, x_o) < 1e-8
    # the hierarchy is the same: dims and directions exactly, weights to rounding
    lv_o = O.mg_levels(nx, ny, nz, vals)
    lv_g = sol.levels()
    assert [(tuple(int(t) for t in a[0]), a[1]) for a in lv_o] == [(a[0], a[1]) for a in lv_g]
    for a, g in zip(lv_o, lv_g):
        assert abs(a[2] - g[2]) <= 1e-14 * abs(a[2])


def test_pfmg_level_operators_match_oracle(torch_cuda):
    S, (nx, ny, nz), vals = _box(SMALL, "p")
    box = F.StructBox(nx, ny, nz, vals)
    sol = F.Solver(box, tol=1e-8, maxit=5, kdim=5, precond="pfmg")
    lv = sol.levels()
    assert len(lv) == len(O.mg_levels(nx, ny, nz, vals))
    for l in range(len(lv)):
        want = O.mg_level_matrix(nx, ny, nz, vals, l).reshape(-1, 7)
        got = sol.level_matrix(l).reshape(-1, 7)
        scale = np.max(np.abs(want), axis=1, keepdims=True) + 1e-300
        assert np.max(np.abs(got - want) / scale) < 1e-13, l


def test_pfmg_vcycle_on_poisson_matches_oracle(torch_cuda):
    from test_oracle_pins import _poisson
    nx, ny, nz = 40, 24, 20
    vals = _poisson(nx, ny, nz, 1.0, 1.1, 0.9)
    b = np.random.default_rng(12).standard_normal(nx * ny * nz)
    box = F.StructBox(nx, ny, nz, vals)
    for relax in (1, 2):
        want = O.precond_apply(nx, ny, nz, vals, b, precond="pfmg", relax_type=relax)
        got = F.pfmg_apply(box, b, relax_type=relax)
        assert rel_l2(got, want) < 1e-12, relax


@pytest.mark.parametrize("relax", [0, 1, 2, 3])
def test_pfmg_vcycle_matches_oracle(torch_cuda, relax):
    S, (nx, ny, nz), vals = _box(SMALL, "p")
    rng = np.random.default_rng(11)
    b = rng.standard_normal(nx * ny * nz)
    want = O.precond_apply(nx, ny, nz, vals, b, precond="pfmg", relax_type=relax)
    box = F.StructBox(nx, ny, nz, vals)
    got = F.pfmg_apply(box, b, relax_type=relax)
    assert rel_l2(got, want) < 1e-12


@pytest.mark.parametrize("coarse", [200, 1000])
def test_pfmg_coarse_size_matches_oracle(torch_cuda, coarse):
    S, (nx, ny, nz), vals = _box(SMALL, "p")
    b = np.random.default_rng(13).standard_normal(nx * ny * nz)
    try:
        O.set_coarse_size(coarse)
        want = O.precond_apply(nx, ny, nz, vals, b, precond="pfmg")
        lv_o = O.mg_levels(nx, ny, nz, vals)
    finally:
        O.set_coarse_size(0)
    box = F.StructBox(nx, ny, nz, vals)
    got = F.pfmg_apply(box, b, coarse_size=coarse)
    assert rel_l2(got, want) < 1e-12
    sol = F.Solver(box, tol=1e-8, maxit=5, kdim=5, precond="pfmg", coarse_size=coarse)
    assert [(tuple(int(t) for t in a[0]), a[1]) for a in lv_o] == [(a[0], a[1]) for a in sol.levels()]
    assert int(np.prod(lv_o[-1][0])) <= coarse


def test_pfmg_fp32_vcycle_close_to_fp64(torch_cuda):
    S, (nx, ny, nz), vals = _box(SMALL, "p")
    b = np.random.default_rng(2).standard_normal(nx * ny * nz)
    box = F.StructBox(nx, ny, nz, vals)
    x64 = F.pfmg_apply(box, b, mg_bits=64)
    x32 = F.pfmg_apply(box, b, mg_bits=32)
    assert rel_l2(x32, x64) < 1e-4


def test_flexible_gmres_fp32_mg_reaches_fp64_tolerance(torch_cuda):
    S, (nx, ny, nz), vals = _box(SMALL, "p")
    fl = S.get("fluid_p").astype(bool)
    b = np.random.default_rng(9).standard_normal(nx * ny * nz)
    b[~fl] = 0.0
    box = F.StructBox(nx, ny, nz, vals)
    sol = F.Solver(box, tol=1e-10, maxit=200, kdim=30, precond="pfmg", flexible=True, mg_bits=32)
    x, its, rel, rc = sol.solve(b)
    assert rc == 0 and rel <= 1e-10
    A = box_csr(nx, ny, nz, vals)
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-9


def test_zero_rhs_gives_zero_iterations(torch_cuda):
    S, (nx, ny, nz), vals = _box(TINY, "v")
    box = F.StructBox(nx, ny, nz, vals)
    sol = F.Solver(box, tol=1e-8, maxit=50, kdim=10, precond="jacobi")
    x, its, rel, rc = sol.solve(np.zeros(nx * ny * nz))
    assert rc == 0 and its == 0 and not np.any(x)


def test_maxit_returns_conv_flag(torch_cuda):
    S, (nx, ny, nz), vals = _box(TINY, "p")
    fl = S.get("fluid_p").astype(bool)
    b = np.random.default_rng(1).standard_normal(nx * ny * nz)
    b[~fl] = 0.0
    box = F.StructBox(nx, ny, nz, vals)
    sol = F.Solver(box, tol=1e-14, maxit=2, kdim=10, precond="jacobi")
    x, its, rel, rc = sol.solve(b)
    assert rc == F.ERROR_CONV and its == 2 and rel > 1e-14



def test_one_point_workgroup_size_changes_no_bit(torch_cuda):
    """The one-point multigrid kernels of the small levels on workgroups of 64, 128 or
    256 threads ("mg_block"): one V-cycle and a GMRES + PFMG solve of the pressure
    operator give the same bits."""
    kw = dict(SMALL)
    S, dims, vals = _box(kw, "p")
    b = np.random.default_rng(71).standard_normal(int(np.prod(dims)))
    out = []
    try:
        for blk in (64, 128, 256):
            F.set_tuning("mg_block", blk)
            box = F.StructBox(*dims, vals)
            y = F.pfmg_apply(box, b)
            x, its, _, rc = F.Solver(box, tol=1e-10, maxit=200, kdim=30, precond="pfmg").solve(b)
            assert rc == 0
            out.append((y, x, its))
    finally:
        F.set_tuning("mg_block", 64)
    for y, x, its in out[1:]:
        assert bits_equal(y, out[0][0]) and bits_equal(x, out[0][1]) and its == out[0][2]

"""GPU parity of the hot copies of z-extruded operators.

An operator whose z-planes away from the walls repeat (the channel's operators:
a full-height cylinder between z walls) is solved from a copy holding m planes
per wall and one interior plane; points whose row differs from the interior
plane (the pressure operator's assembly rounds differently on a few rows) are
flagged and read their own row.  Checks:

* the reference setup's four operators are detected as extruded (the pressure
  operator with flagged points); an operator that varies at every point and a
  box of nz < 5 are not;
* on extruded operators, a matvec, one V-cycle and GMRES (Jacobi, PFMG) match
  the oracle to the Struct tests' tolerances (1e-13 matvec, 1e-12 V-cycle,
  1e-9 solutions at rtol 1e-11);
* the same calls give the same bits with the copy switched off ("z_extrude" 0),
  on the z-marching kernels and on the one-point ones, FP64 and FP32 (V-cycle
  and Jacobi), first coarsening in z (odd and even nz), y and x, with and
  without flagged points;
* a simulation's steps are bitwise the same either way.
"""
import numpy as np
import pytest

import fpc_amd as F
import oracle as O
from helpers import CFG1, SMALL, bits_equal, box_csr, rel_l2
from test_gpu_tiles import _aniso

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def restore_tuning(torch_cuda):
    yield
    F.set_tuning("z_extrude", 1)
    F.set_tuning("tile_min_points", 1 << 19)
    F.set_tuning("stencil_kz", 0)
    F.set_tuning("mg_wide_finest", 1)


def _extruded(nx, ny, nz, scale, seed, flagged=0.0):
    """A random nonsymmetric M-matrix in x and y, repeated along z: plane 0, planes
    1..nz-2 and plane nz-1 from the three planes of a 3-plane operator; then a
    fraction `flagged` of the points gets one coefficient moved by one ulp."""
    base = _aniso(nx, ny, 3, scale, seed).reshape(3, ny, nx, 7)
    v = np.empty((nz, ny, nx, 7))
    v[0], v[1:-1], v[-1] = base[0], base[1], base[2]
    v = v.reshape(-1, 7)
    rng = np.random.default_rng(seed + 100)
    for t in rng.choice(len(v), int(flagged * len(v)), replace=False):
        s = int(rng.integers(7))
        v[t, s] = np.nextafter(v[t, s], np.inf)
    return v.reshape(-1)


CASES = [  # dims, per-direction coupling scale, expected first coarsening direction
    ((96, 40, 37), (1.0, 1.0, 4.0), 2),
    ((80, 24, 36), (1.0, 1.0, 4.0), 2),
    ((96, 40, 37), (1.0, 4.0, 1.0), 1),
    ((96, 40, 21), (4.0, 1.0, 1.0), 0),
    ((64, 24, 100), (1.0, 1.0, 4.0), 2),  # two flag words per column (planes 0-63, 64-99)
]
IDS = ["z-odd", "z-even", "y", "x", "z-deep"]


def test_reference_operators_are_extruded():
    S = O.Setup(O.default_params(**SMALL))
    nx, ny, nz = SMALL["nx"], SMALL["ny"], SMALL["nz"]
    dims = {"Au_vals": (nx - 1, ny, nz), "Av_vals": (nx, ny - 1, nz), "Aw_vals": (nx, ny, nz - 1),
            "Ap_vals": (nx, ny, nz)}
    for name, d in dims.items():
        m, flagged = F.StructBox(*d, S.get(name)).extruded()
        assert m >= 1, name
        assert (flagged > 0) == (name == "Ap_vals"), (name, flagged)
    assert F.StructBox(nx, ny, nz, _aniso(nx, ny, nz, (1, 1, 1), 5)).extruded() == (0, 0)
    assert F.StructBox(nx, ny, 4, _extruded(nx, ny, 4, (1, 1, 1), 0)).extruded() == (0, 0)


def _calls(box, b, mg_bits=64):
    """One V-cycle, GMRES + PFMG, GMRES + Jacobi (FP64 and FP32 1/diag)."""
    flex = mg_bits == 32
    y = F.pfmg_apply(box, b, mg_bits=mg_bits)
    sp = F.Solver(box, tol=1e-11, maxit=300, kdim=30, precond="pfmg", flexible=flex, mg_bits=mg_bits)
    xp, itp, _, rcp = sp.solve(b)
    xj, itj, _, rcj = F.Solver(box, tol=1e-11, maxit=400, kdim=30, precond="jacobi").solve(b)
    xj32, itj32, _, _ = F.Solver(box, tol=1e-11, maxit=400, kdim=30, precond="jacobi", jacobi_bits=32).solve(b)
    assert rcp == 0 and rcj == 0
    return dict(y=y, xp=xp, itp=itp, xj=xj, itj=itj, xj32=xj32, itj32=itj32, lev=sp.levels_extruded())


@pytest.mark.parametrize("flagged", [0.0, 0.003], ids=["exact", "flagged"])
@pytest.mark.parametrize("tiles", [True, False], ids=["tiled", "one-point"])
@pytest.mark.parametrize("dims,scale,cdir", CASES, ids=IDS)
def test_extruded_matches_oracle_and_full_storage(dims, scale, cdir, tiles, flagged):
    nx, ny, nz = dims
    F.set_tuning("tile_min_points", 0 if tiles else 1 << 40)
    vals = _extruded(nx, ny, nz, scale, seed=nz + 3, flagged=flagged)
    rng = np.random.default_rng(nz)
    b = rng.standard_normal(nx * ny * nz)
    x = rng.standard_normal(nx * ny * nz)
    assert O.mg_levels(nx, ny, nz, vals)[0][1] == cdir
    box = F.StructBox(nx, ny, nz, vals)
    m, nflag = box.extruded()
    assert m >= 1 and (nflag > 0) == (flagged > 0), (m, nflag)
    ext = _calls(box, b)
    assert ext["lev"][0][0] >= 1, ext["lev"]
    # against the oracle
    assert rel_l2(box.matvec(x), O.struct_matvec(nx, ny, nz, vals, x)) < 1e-13
    assert rel_l2(ext["y"], O.precond_apply(nx, ny, nz, vals, b, precond="pfmg")) < 1e-12
    A = box_csr(nx, ny, nz, vals)
    for key, pc in (("xp", "pfmg"), ("xj", "jacobi")):
        x_o, its_o, _, rc_o = O.gmres(nx, ny, nz, vals, b, tol=1e-11, maxit=400, kdim=30, precond=pc)
        assert rc_o == 0
        assert abs(ext["itp" if pc == "pfmg" else "itj"] - its_o) <= 1, pc
        assert rel_l2(ext[key], x_o) < 1e-9, pc
        assert np.linalg.norm(b - A @ ext[key]) / np.linalg.norm(b) < 2e-11, pc
    # against the full planes: the same bits
    F.set_tuning("z_extrude", 0)
    box_f = F.StructBox(nx, ny, nz, vals)
    assert box_f.extruded() == (0, 0)
    full = _calls(box_f, b)
    assert all(lv == (0, 0) for lv in full["lev"])
    for key in ("y", "xp", "xj", "xj32"):
        assert bits_equal(ext[key], full[key]), key
    for key in ("itp", "itj", "itj32"):
        assert ext[key] == full[key], key


@pytest.mark.parametrize("kz", [6, 10, 100])
def test_flag_words_across_chunks(kz):
    """Flags of planes 0-63 and 64-99 in two words per column, with z-chunks of the
    marching tiles that cross plane 64 inside their loop (a chunk length above 62 is
    clamped to 62, so a chunk's planes still span at most two words): the same bits
    as the full planes, run after run."""
    nx, ny, nz = 64, 24, 100
    F.set_tuning("tile_min_points", 0)
    F.set_tuning("stencil_kz", kz)
    vals = _extruded(nx, ny, nz, (1.0, 1.0, 4.0), seed=41, flagged=0.01)
    b = np.random.default_rng(42).standard_normal(nx * ny * nz)
    box = F.StructBox(nx, ny, nz, vals)
    assert box.extruded()[1] > 0
    ext = [_calls(box, b) for _ in range(2)]
    F.set_tuning("z_extrude", 0)
    full = _calls(F.StructBox(nx, ny, nz, vals), b)
    for e in ext:
        for key in ("y", "xp", "xj", "xj32"):
            assert bits_equal(e[key], full[key]), key


@pytest.mark.parametrize("dims,scale,cdir", [CASES[0], CASES[3]], ids=["z-odd", "x"])
def test_extruded_fp32_vcycle_equals_full_storage(dims, scale, cdir):
    nx, ny, nz = dims
    F.set_tuning("tile_min_points", 0)
    vals = _extruded(nx, ny, nz, scale, seed=nz + 5, flagged=0.003)
    b = np.random.default_rng(3).standard_normal(nx * ny * nz)
    ext = _calls(F.StructBox(nx, ny, nz, vals), b, mg_bits=32)
    F.set_tuning("z_extrude", 0)
    full = _calls(F.StructBox(nx, ny, nz, vals), b, mg_bits=32)
    assert bits_equal(ext["y"], full["y"])
    assert ext["itp"] == full["itp"] and bits_equal(ext["xp"], full["xp"])
    assert rel_l2(ext["y"], F.pfmg_apply(F.StructBox(nx, ny, nz, vals), b)) < 1e-4


@pytest.mark.parametrize("tiles", [True, False], ids=["tiled", "one-point"])
@pytest.mark.parametrize("dims,scale,cdir", [CASES[0], CASES[2], CASES[3]], ids=["z", "y", "x"])
def test_fp32_vcycle_reads_and_writes_fp64_on_the_finest_level(dims, scale, cdir, tiles):
    """The FP32 V-cycle's finest-level kernels read the FP64 rhs and write the FP64
    result directly (rounding as they read and write): the same bits as converting
    to and from FP32 copies with separate kernels, on the z-restricting, tiled and
    one-point finest-level kernels, for one V-cycle and a flexible GMRES solve."""
    nx, ny, nz = dims
    F.set_tuning("tile_min_points", 0 if tiles else 1 << 40)
    vals = _extruded(nx, ny, nz, scale, seed=nz + 11, flagged=0.003)
    b = np.random.default_rng(nz + 1).standard_normal(nx * ny * nz)
    out = []
    for wide in (1, 0):
        F.set_tuning("mg_wide_finest", wide)
        box = F.StructBox(nx, ny, nz, vals)
        y = F.pfmg_apply(box, b, mg_bits=32)
        sp = F.Solver(box, tol=1e-11, maxit=300, kdim=30, precond="pfmg", flexible=True, mg_bits=32)
        x, its, _, rc = sp.solve(b)
        assert rc == 0
        out.append((y, x, its))
    assert bits_equal(out[0][0], out[1][0]) and bits_equal(out[0][1], out[1][1]) and out[0][2] == out[1][2]
    assert np.array_equal(out[0][0], out[0][0].astype(np.float32).astype(np.float64))  # FP32 values
    assert rel_l2(out[0][0], F.pfmg_apply(F.StructBox(nx, ny, nz, vals), b)) < 1e-4


def test_simulation_steps_equal_full_storage():
    """128x32x32 with the z-marching kernels on every box of at least 32768 points: with
    the copies on, the tiled matvecs and V-cycle levels read the hot copies (the pressure
    operator's with flagged points); with them off, the full planes.  Same bits."""
    kw = dict(CFG1, nsteps=4, init_profile=1, noise_amp=1e-3 * 0.45, seed=2)
    F.set_tuning("tile_min_points", 0)
    out = []
    for flag in (1, 0):
        F.set_tuning("z_extrude", flag)
        sim = F.Sim(F.default_params(**kw))
        sim.step(4)
        ext = {c: (sim.scalar("zext_m_" + c), sim.scalar("zext_flagged_" + c)) for c in "uvwp"}
        if flag:
            assert all(m >= 1 for m, _ in ext.values()), ext
            assert ext["p"][1] > 0, ext
        else:
            assert all(m == 0 for m, _ in ext.values()), ext
        out.append({k: sim.get(k) for k in ("u", "v", "w", "p")})
    for k in ("u", "v", "w", "p"):
        assert bits_equal(out[0][k], out[1][k]), k


def test_new_values_replace_the_hot_copies():
    """Values set again after assembly drop the copies; the next solve plans them anew
    (an extruded operator replaced by one that is not, and back) and matches a fresh box."""
    nx, ny, nz = 96, 40, 21
    F.set_tuning("tile_min_points", 0)
    a1 = _extruded(nx, ny, nz, (1.0, 1.0, 4.0), seed=31, flagged=0.002)
    a2 = _aniso(nx, ny, nz, (1.0, 1.0, 4.0), seed=32)
    b = np.random.default_rng(33).standard_normal(nx * ny * nz)
    box = F.StructBox(nx, ny, nz, a1)
    sol = F.Solver(box, tol=1e-11, maxit=300, kdim=30, precond="pfmg")
    sol.solve(b)
    for vals, extruded in ((a2, False), (a1, True)):
        vals = np.ascontiguousarray(vals)
        F.check(F.lib().fpc_SetMatrix3D(box.A, nx, ny, nz, vals.ctypes.data), "SetMatrix3D")
        x, its, _, rc = sol.solve(b)
        fresh = F.StructBox(nx, ny, nz, vals)
        x_f, its_f, _, rc_f = F.Solver(fresh, tol=1e-11, maxit=300, kdim=30, precond="pfmg").solve(b)
        assert rc == 0 and rc_f == 0 and its == its_f
        assert bits_equal(x, x_f)
        assert box.extruded() == fresh.extruded()
        assert (box.extruded()[0] > 0) == extruded

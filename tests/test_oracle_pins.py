"""Pins of the CPU oracle (no GPU).

The reference ships no tests or golden vectors, so every oracle function is
checked against what the mathematics fixes, computed here independently:
closed-form geometry and masks, interior and wall stencils, Ap = -D G with
scipy, exact solves with scipy.sparse.linalg.spsolve, the Jacobi closed form,
a numpy two-level cycle of the stated multigrid definition, exactness of the
centred convection on linear fields, and the reference quirks listed in
main.cu / gpu_dense_ops.hpp / immersed_operators.hpp.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import oracle as O
from helpers import CFG1, SMALL, TINY, box_csr, csr_1based, rel_l2


def P(kw):
    return O.default_params(**kw)


@pytest.fixture(scope="module")
def small():
    return O.Setup(P(SMALL)), SMALL


@pytest.fixture(scope="module")
def tiny():
    return O.Setup(P(TINY)), TINY


def full(kw, key):
    """Full parameter dict with the reference defaults (common.hpp:45-85) under the overrides."""
    d = {k: getattr(O.default_params(), k) for k, _ in O.PARAM_FIELDS}
    d.update(kw)
    return d[key]


# ---------------------------------------------------------------- geometry (immersed_geometry.hpp)
def test_beta_independent_subsampling(small):
    S, kw = small
    nx, ny = kw["nx"], kw["ny"]
    dx, dy = full(kw, "Lx") / nx, full(kw, "Ly") / ny
    xc, yc, R = full(kw, "xcirc"), full(kw, "ycirc"), full(kw, "Rcirc")
    sub = (np.arange(10) + 0.5) / 10.0
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")  # (ny, nx)
    xs = (i * dx)[..., None, None] + sub[:, None] * dx
    ys = (j * dy)[..., None, None] + sub[None, :] * dy
    out = ((xs - xc) * (xs - xc) + (ys - yc) * (ys - yc)) > R * R
    beta = out.sum(axis=(2, 3)) / 100.0
    assert np.array_equal(S.get("beta2d").reshape(ny, nx), beta)
    # the blocked area approximates the circle's area
    blocked = ((1.0 - beta) * dx * dy).sum()
    assert abs(blocked - np.pi * R * R) < 0.05 * np.pi * R * R


def test_alpha_chord_formula(small):
    S, kw = small
    nx, ny = kw["nx"], kw["ny"]
    dx, dy = full(kw, "Lx") / nx, full(kw, "Ly") / ny
    xc, yc, R = full(kw, "xcirc"), full(kw, "ycirc"), full(kw, "Rcirc")
    au = S.get("alpha_u2d").reshape(ny, nx + 1)
    for i in range(nx + 1):
        x = i * dx
        for j in range(ny):
            want = 1.0
            if abs(x - xc) < R:
                h = np.sqrt(max(0.0, R * R - (x - xc) ** 2))
                ov = max(0.0, min((j + 1) * dy, yc + h) - max(j * dy, yc - h))
                want = max(0.0, min(1.0, 1.0 - ov / dy))
            assert au[j, i] == want
    # faces away from the cylinder are fully open
    assert np.all(au[:, : int((xc - R) / dx) - 1] == 1.0)


def test_masks_closed_form_and_z_invariance(small):
    S, kw = small
    nx, ny, nz = kw["nx"], kw["ny"], kw["nz"]
    dx, dy = full(kw, "Lx") / nx, full(kw, "Ly") / ny
    xc, yc, R = full(kw, "xcirc"), full(kw, "ycirc"), full(kw, "Rcirc")
    fl = S.get("fluid2d").reshape(ny, nx).astype(bool)
    beta = S.get("beta2d").reshape(ny, nx)
    assert np.array_equal(fl, beta > 1e-8) or S.count("n_pruned") > 0
    au = S.get("active_u").reshape(nz, ny, nx + 1).astype(bool)
    for k in range(nz):
        assert np.array_equal(au[k], au[0])
    # u face i (1..nx-1) active iff its centre is outside the circle and a neighbour cell is fluid
    jj, ii = np.meshgrid(np.arange(ny), np.arange(1, nx), indexing="ij")
    x, y = ii * dx, (jj + 0.5) * dy
    want = (((x - xc) * (x - xc) + (y - yc) * (y - yc)) > R * R) & (fl[jj, ii - 1] | fl[jj, ii])
    assert np.array_equal(au[0][:, 1:nx], want)
    assert not au[0][:, 0].any() and not au[0][:, nx].any()
    aqu = S.get("active_qu").reshape(nz, ny, nx + 1).astype(bool)
    assert np.array_equal(aqu[0][:, nx], fl[:, nx - 1])  # outflow faces of fluid cells
    aw = S.get("active_w").reshape(nz + 1, ny, nx).astype(bool)
    for k in range(nz + 1):  # w faces, wall faces included, follow the fluid column
        assert np.array_equal(aw[k], fl)


def test_ids_are_one_based_cumulative(small):
    S, _ = small
    for m, i in (("active_u", "id_u"), ("fluid_p", "id_p"), ("active_qw", "id_qw")):
        mask = S.get(m).astype(bool)
        ids = S.get(i)
        assert np.array_equal(ids, np.where(mask, np.cumsum(mask), 0))
    assert S.count("Np") == int(S.get("fluid_p").sum())


def test_pruning_removes_isolated_columns():
    # a cylinder nearly touching the wall leaves columns with no open face
    kw = dict(nx=40, ny=16, nz=4, Lx=1.0, xcirc=0.5, ycirc=0.058, Rcirc=0.06)
    S = O.Setup(P(kw))
    assert S.count("n_pruned") > 0
    beta = S.get("beta2d")
    fl = S.get("fluid2d").astype(bool)
    assert np.all(beta[fl] > 1e-8)
    assert int(((beta > 1e-8) & ~fl).sum()) == S.count("n_pruned")


# ---------------------------------------------------------------- assembly (immersed_operators.hpp)
def _interior_scaled(kw):
    nx, ny, nz = kw["nx"], kw["ny"], kw["nz"]
    dx, dy, dz = full(kw, "Lx") / nx, full(kw, "Ly") / ny, full(kw, "Lz") / nz
    nu, rho, dt = full(kw, "nu"), full(kw, "rho"), full(kw, "dt")
    cx, cy, cz = nu / dx ** 2, nu / dy ** 2, nu / dz ** 2
    diag = rho / dt + 2 * (cx + cy + cz)
    s = diag + 2 * (cx + cy + cz)
    return np.array([diag, -cx, -cx, -cy, -cy, -cz, -cz]) / s


def test_interior_momentum_stencil_closed_form(small):
    S, kw = small
    nx, ny, nz = kw["nx"], kw["ny"], kw["nz"]
    A = S.get("Au_vals").reshape(nz, ny, nx - 1, 7)
    row = A[nz // 2, 2, 3]  # far from the cylinder and the walls
    assert np.allclose(row, _interior_scaled(kw), rtol=4e-16, atol=0)
    # south wall row: closure dy / (0.5 dy) = 2 (immersed_operators.hpp:41)
    wall = A[nz // 2, 0, 3]
    assert wall[3] == 0.0 and wall[4] < 0.0
    nu, dy = full(kw, "nu"), full(kw, "Ly") / ny
    cy = nu / dy ** 2
    diag_unscaled = -wall[0] / wall[4] * cy
    dx, dz = full(kw, "Lx") / nx, full(kw, "Lz") / nz
    want = full(kw, "rho") / full(kw, "dt") + 2 * nu / dx ** 2 + cy + 2 * cy + 2 * nu / dz ** 2
    assert abs(diag_unscaled - want) < 1e-12 * want


def test_inactive_rows_are_identity(small):
    S, kw = small
    for name, act_name, shape, sl in (
        ("Au_vals", "active_u", (kw["nz"], kw["ny"], kw["nx"] + 1), np.s_[:, :, 1:-1]),
        ("Ap_vals", "fluid_p", (kw["nz"], kw["ny"], kw["nx"]), np.s_[:, :, :]),
    ):
        A = S.get(name).reshape(-1, 7)
        act = S.get(act_name).reshape(shape)[sl].reshape(-1).astype(bool)
        assert np.all(A[~act, 0] == 1.0) and np.all(A[~act, 1:] == 0.0)


def test_pressure_matrix_is_minus_DG(small):
    S, kw = small
    nx, ny, nz = kw["nx"], kw["ny"], kw["nz"]
    Nq = S.count("Nqu") + S.count("Nqv") + S.count("Nqw")
    Np = S.count("Np")
    D = csr_1based(S.get("D_rowptr"), S.get("D_colind"), S.get("D_vals"), Nq)
    G = csr_1based(S.get("G_rowptr"), S.get("G_colind"), S.get("G_vals"), Np)
    Ap = (-(D @ G)).toarray()
    box = box_csr(nx, ny, nz, S.get("Ap_vals")).toarray()
    fl = S.get("fluid_p").astype(bool)
    sub = box[np.ix_(fl, fl)]
    assert np.allclose(sub, Ap, rtol=1e-14, atol=1e-9)
    # off-diagonal entries are single products: exact
    off = ~np.eye(Np, dtype=bool)
    assert np.array_equal(sub[off], Ap[off])
    # solid rows are decoupled identities
    assert np.all(np.diag(box)[~fl] == 1.0)


def test_gradient_and_divergence_rows(small):
    S, kw = small
    nx, dx = kw["nx"], full(kw, "Lx") / kw["nx"]
    Np = S.count("Np")
    G = csr_1based(S.get("G_rowptr"), S.get("G_colind"), S.get("G_vals"), Np).toarray()
    nqu = S.count("Nqu")
    g = G[:nqu]
    assert set(np.unique(np.round(g[g != 0] * dx, 12))) <= {-2.0, -1.0, 1.0}
    # outflow rows carry -2/dx (immersed_operators.hpp:260-266)
    assert np.sum(np.isclose(g * dx, -2.0)) == int(S.get("fluid2d").reshape(kw["ny"], nx)[:, -1].sum()) * kw["nz"]
    # divergence of a uniform x-flow u = 1 in an open cell is 0
    D = csr_1based(S.get("D_rowptr"), S.get("D_colind"), S.get("D_vals"), S.count("Nqu") + S.count("Nqv") + S.count("Nqw"))
    q = np.zeros(D.shape[1])
    q[:nqu] = 1.0
    dq = D @ q
    beta = S.get("beta_p") if False else None
    fl = S.get("fluid_p").astype(bool)
    idx = np.where(fl)[0]
    nxy = nx * kw["ny"]
    i = idx % nx
    interior = (i > 0) & (i < nx - 1)
    alpha = S.get("alpha_u2d")
    # cells whose two x-faces are fully open see zero divergence
    j = (idx // nx) % kw["ny"]
    open2 = (alpha[i + j * (nx + 1)] == 1.0) & (alpha[i + 1 + j * (nx + 1)] == 1.0)
    assert np.max(np.abs(dq[interior & open2])) < 1e-9


def test_reference_quirks(small):
    S, kw = small
    nx, ny, nz = kw["nx"], kw["ny"], kw["nz"]
    # Q1: only u has a Dirichlet rhs, unit amplitude, on the first interior face
    assert not S.get("rhs_v_bc").any() and not S.get("rhs_w_bc").any()
    bc = S.get("rhs_u_bc")
    ids = S.get("id_u").reshape(nz, ny, nx + 1)
    first = ids[:, :, 1][ids[:, :, 1] > 0]
    others = np.setdiff1d(np.arange(1, len(bc)), first)
    assert np.all(bc[others] == 0.0) and np.all(bc[first] > 0.0)
    nu, dx = full(kw, "nu"), full(kw, "Lx") / nx
    dy, dz, H = full(kw, "Ly") / ny, full(kw, "Lz") / nz, full(kw, "Ly")
    k, j = nz // 2, ny // 2
    y, z = (j + 0.5) * dy, (k + 0.5) * dz
    prof = 16.0 * 1.0 * y * z * (H - y) * (H - z) / H ** 4
    assert abs(bc[ids[k, j, 1]] - nu / dx ** 2 * prof) < 1e-14 * abs(bc[ids[k, j, 1]])
    # divergence bc uses Um and lives on the i = 0 cells only
    dbc = S.get("div_bc")
    idp = S.get("id_p").reshape(nz, ny, nx)
    assert np.all(dbc[idp[:, :, 1:][idp[:, :, 1:] > 0]] == 0.0)
    assert np.all(dbc[idp[:, :, 0][idp[:, :, 0] > 0]] < 0.0)


# ---------------------------------------------------------------- solvers
def test_struct_matvec_equals_scipy(small):
    S, kw = small
    nx, ny, nz = kw["nx"], kw["ny"], kw["nz"]
    vals = S.get("Ap_vals")
    x = np.random.default_rng(0).standard_normal(nx * ny * nz)
    assert rel_l2(O.struct_matvec(nx, ny, nz, vals, x), box_csr(nx, ny, nz, vals) @ x) < 1e-15


@pytest.mark.parametrize("which,precond", [("u", "jacobi"), ("w", "jacobi"), ("u", "none"), ("p", "pfmg")])
def test_gmres_solution_equals_spsolve(tiny, which, precond):
    S, kw = tiny
    nx, ny, nz = kw["nx"], kw["ny"], kw["nz"]
    dims = {"u": (nx - 1, ny, nz), "w": (nx, ny, nz - 1), "p": (nx, ny, nz)}[which]
    vals = S.get({"u": "Au_vals", "w": "Aw_vals", "p": "Ap_vals"}[which])
    A = box_csr(*dims, vals).tocsc()
    b = np.random.default_rng(1).standard_normal(A.shape[0])
    x, its, rel, rc = O.gmres(*dims, vals, b, tol=1e-12, maxit=2000, kdim=50, precond=precond)
    assert rc == 0 and rel <= 1e-12
    xs = spl.spsolve(A, b)
    assert rel_l2(x, xs) < 1e-9
    assert np.linalg.norm(b - A @ x) <= 1.0001e-12 * np.linalg.norm(b)


def test_jacobi_one_sweep_is_division(tiny):
    S, kw = tiny
    nx, ny, nz = kw["nx"] - 1, kw["ny"], kw["nz"]
    vals = S.get("Au_vals")
    b = np.random.default_rng(2).standard_normal(nx * ny * nz)
    x = O.precond_apply(nx, ny, nz, vals, b, precond="jacobi")
    assert np.array_equal(x, b / vals.reshape(-1, 7)[:, 0])


def test_gmres_zero_rhs_and_maxit(tiny):
    S, kw = tiny
    nx, ny, nz = kw["nx"], kw["ny"], kw["nz"]
    vals = S.get("Ap_vals")
    x, its, rel, rc = O.gmres(nx, ny, nz, vals, np.zeros(nx * ny * nz), tol=1e-8, maxit=10, kdim=5)
    assert its == 0 and rc == 0 and not x.any()
    b = np.random.default_rng(3).standard_normal(nx * ny * nz)
    x, its, rel, rc = O.gmres(nx, ny, nz, vals, b, tol=1e-15, maxit=3, kdim=5)
    assert rc == 256 and its == 3


# ---------------------------------------------------------------- multigrid definition
def _poisson(nx, ny, nz, hx=1.0, hy=1.0, hz=1.0):
    v = np.zeros((nz, ny, nx, 7))
    cx, cy, cz = 1 / hx ** 2, 1 / hy ** 2, 1 / hz ** 2
    v[..., 0] = 2 * (cx + cy + cz)
    v[..., 1] = v[..., 2] = -cx
    v[..., 3] = v[..., 4] = -cy
    v[..., 5] = v[..., 6] = -cz
    v[:, :, 0, 1] = 0; v[:, :, -1, 2] = 0; v[:, 0, :, 3] = 0; v[:, -1, :, 4] = 0; v[0, :, :, 5] = 0; v[-1, :, :, 6] = 0
    return v.reshape(-1)


def _two_level_numpy(nx, ny, nz, vals, d, omega, b):
    """Independent numpy two-level V(1,1) of the stated definition: operator-collapse
    interpolation, R = P^T; coarse couplings across d = Galerkin couplings summed over
    the offsets along d, couplings along d = P^T A~ P with A~ the stencil collapsed
    across d, centre keeping the Galerkin row sum; weighted Jacobi; exact coarse solve."""
    A = box_csr(nx, ny, nz, vals).toarray()
    n = [nx, ny, nz]
    nc = list(n)
    nc[d] = n[d] // 2
    N, Nc = nx * ny * nz, nc[0] * nc[1] * nc[2]
    fi = lambda p: (p[2] * ny + p[1]) * nx + p[0]
    ci = lambda p: (p[2] * nc[1] + p[1]) * nc[0] + p[0]
    cpos = lambda r: [r % nc[0], (r // nc[0]) % nc[1], r // (nc[0] * nc[1])]
    Pm = np.zeros((N, Nc))
    At = A.copy()  # stencil collapsed across d
    for r in range(N):
        p = [r % nx, (r // nx) % ny, r // (nx * ny)]
        for e in range(3):
            if e == d:
                continue
            for sg in (-1, 1):
                q = list(p); q[e] += sg
                if 0 <= q[e] < n[e]:
                    At[r, r] += A[r, fi(q)]
                    At[r, fi(q)] = 0.0
    for r in range(N):
        p = [r % nx, (r // nx) % ny, r // (nx * ny)]
        if p[d] % 2 == 1:
            q = list(p); q[d] = (p[d] - 1) // 2; Pm[r, ci(q)] = 1.0
            continue
        ctr = At[r, r]
        if p[d] - 1 >= 0:
            q = list(p); q[d] -= 1
            qc = list(p); qc[d] = p[d] // 2 - 1; Pm[r, ci(qc)] = -A[r, fi(q)] / ctr
        if p[d] + 1 < n[d]:
            q = list(p); q[d] += 1
            qc = list(p); qc[d] = p[d] // 2; Pm[r, ci(qc)] = -A[r, fi(q)] / ctr
    Ag = Pm.T @ A @ Pm
    Ad = Pm.T @ At @ Pm
    L = np.zeros((Nc, Nc))
    for r in range(Nc):
        p = cpos(r)
        for c in np.nonzero(Ag[r])[0]:
            q = cpos(c)
            off = [q[t] - p[t] for t in range(3)]
            cross = [t for t in range(3) if t != d and off[t] != 0]
            if cross:
                q2 = list(q); q2[d] = p[d]; L[r, ci(q2)] += Ag[r, c]
        for c in np.nonzero(Ad[r])[0]:
            if c != r:
                L[r, c] += Ad[r, c]
        L[r, r] = Ag[r].sum() - (L[r].sum() - L[r, r])
    Dg = np.diag(A)
    x = omega * b / Dg
    r_ = b - A @ x
    x = x + Pm @ np.linalg.solve(L, Pm.T @ r_)
    return x + omega * (b - A @ x) / Dg


# fine levels of 64 < N <= 128 points: exactly one coarsening, then the coarsest
# level (<= 64 points) is solved exactly
@pytest.mark.parametrize("dims,h", [((8, 4, 4), (1, 1, 1)), ((4, 8, 4), (1, 0.8, 1.2)), ((4, 4, 8), (1.3, 1, 0.9))])
def test_two_level_cycle_matches_numpy_definition(dims, h):
    nx, ny, nz = dims
    vals = _poisson(nx, ny, nz, *h)
    lv = O.mg_levels(nx, ny, nz, vals)
    assert len(lv) == 2 and lv[1][1] == -1
    d, omega = lv[0][1], lv[0][2]
    # direction of strongest coupling and w = 2/(2 + c_d / sum c)
    A = vals.reshape(-1, 7)
    c = [np.abs(A[:, 1 + 2 * e]).sum() + np.abs(A[:, 2 + 2 * e]).sum() for e in range(3)]
    assert d == int(np.argmax(c))
    assert abs(omega - 2.0 / (2.0 + c[d] / sum(c))) < 1e-14  # sums in another order
    b = np.random.default_rng(4).standard_normal(nx * ny * nz)
    want = _two_level_numpy(nx, ny, nz, vals, d, omega, b)
    got = O.precond_apply(nx, ny, nz, vals, b, precond="pfmg", relax_type=1)
    assert rel_l2(got, want) < 1e-12


@pytest.mark.parametrize("dims", [(32, 16, 16), (128, 64, 64)])
def test_vcycle_converges_on_poisson(dims):
    # 128x64x64: the hierarchy stops at <= 256 points (N0 / 2048), not 64
    nx, ny, nz = dims
    vals = _poisson(nx, ny, nz)
    lv = O.mg_levels(nx, ny, nz, vals)
    n_last = int(np.prod(lv[-1][0]))
    cmax = min(1152, max(64, nx * ny * nz // 2048))
    assert n_last <= cmax and int(np.prod(lv[-2][0])) > cmax
    A = box_csr(nx, ny, nz, vals)
    b = np.random.default_rng(6).standard_normal(A.shape[0])
    x = np.zeros_like(b)
    res = []
    for _ in range(6):
        x += O.precond_apply(nx, ny, nz, vals, b - A @ x, precond="pfmg")
        res.append(np.linalg.norm(b - A @ x) / np.linalg.norm(b))
    rates = np.array(res[1:]) / np.array(res[:-1])
    assert np.all(rates < 0.5), res


def test_coarse_size_stops_coarsening_and_solves_exactly():
    # a coarsest size at or above N keeps one level: the V-cycle is the exact solve
    nx, ny, nz = 10, 8, 6
    vals = _poisson(nx, ny, nz, 1.0, 1.2, 0.8)
    b = np.random.default_rng(9).standard_normal(nx * ny * nz)
    try:
        O.set_coarse_size(nx * ny * nz)
        assert len(O.mg_levels(nx, ny, nz, vals)) == 1
        got = O.precond_apply(nx, ny, nz, vals, b, precond="pfmg")
        O.set_coarse_size(200)
        lv = O.mg_levels(nx, ny, nz, vals)
    finally:
        O.set_coarse_size(0)
    want = spl.spsolve(box_csr(nx, ny, nz, vals).tocsc(), b)
    assert rel_l2(got, want) < 1e-13
    # otherwise coarsening stops at the first level with <= n points
    sizes = [int(np.prod(l[0])) for l in lv]
    assert sizes[-1] <= 200 < sizes[-2]
    assert len(O.mg_levels(nx, ny, nz, vals)) > len(lv)


# ---------------------------------------------------------------- dense ops and a whole step
def test_convection_exact_on_linear_and_uniform_fields():
    kw = dict(nx=20, ny=10, nz=8, Lx=1.0, xcirc=5.0, ycirc=5.0, Rcirc=0.01)  # cylinder outside: open channel
    s = O.Sim(P(kw))
    nx, ny, nz = 20, 10, 8
    dx = 1.0 / nx
    x_faces = np.arange(nx + 1) * dx
    u = np.broadcast_to(x_faces, (nz, ny, nx + 1)).copy().reshape(-1)  # u = x, v = w = 0
    s.set("u", u)
    s.set("v", np.zeros(nx * (ny + 1) * nz))
    s.set("w", np.zeros(nx * ny * (nz + 1)))
    s.phase(0)
    s.phase(1)
    cu = s.get("conv_u").reshape(nz, ny, nx + 1)
    uu = s.get("u").reshape(nz, ny, nx + 1)
    inner = np.s_[1:-1, 1:-1, 2:nx - 1]
    assert np.allclose(cu[inner], uu[inner] * 1.0, rtol=1e-12, atol=1e-14)
    s2 = O.Sim(P(kw))
    s2.set("u", np.full((nx + 1) * ny * nz, 0.3))
    s2.set("v", np.zeros(nx * (ny + 1) * nz))
    s2.set("w", np.zeros(nx * ny * (nz + 1)))
    s2.phase(0)
    s2.phase(1)
    cu2 = s2.get("conv_u").reshape(nz, ny, nx + 1)
    assert np.all(cu2[:, :, 2:nx - 1] == 0.0)


def test_projection_makes_the_field_discretely_divergence_free(tiny):
    kw = dict(TINY, init_profile=1, noise_amp=1e-3, seed=5, vel_tol=1e-12, p_tol=1e-13, p_kdim=50)
    s = O.Sim(P(kw))
    s.step(1)
    sc = s.scalars()
    assert sc["relres_p"] <= 1e-13
    assert sc["div_l2"] < 1e-9 * max(1.0, sc["umax"] / (0.41 / 12))


def test_lift_vanishes_for_a_centred_cylinder():
    kw = dict(nx=40, ny=16, nz=6, Lx=1.0, xcirc=0.3, ycirc=0.205, Rcirc=0.06, init_profile=1, vel_tol=1e-12,
              p_tol=1e-12, inlet_ramp_steps=1)
    s = O.Sim(P(kw))
    s.step(2)
    sc = s.scalars()
    assert sc["cd"] > 0.0
    assert abs(sc["cl"]) < 1e-8 * abs(sc["cd"])


def test_seeded_noise_deterministic_and_bounded():
    kw = dict(TINY, init_profile=0, noise_amp=0.01, seed=42)
    a, b = O.Sim(P(kw)), O.Sim(P(kw))
    v = a.get("v")
    assert np.array_equal(v, b.get("v"))
    assert np.max(np.abs(v)) < 0.01 and np.max(np.abs(v)) > 0.009
    assert abs(v[v != 0].mean()) < 0.002
    c = O.Sim(P(dict(kw, seed=43)))
    assert not np.array_equal(v, c.get("v"))


@pytest.mark.slow
def test_cfg1_first_steps_behave():
    s = O.Sim(P(CFG1))
    s.step(2)
    sc = s.scalars()
    assert sc["its_v"] > 0 and sc["relres_u"] <= 1e-8 and sc["relres_p"] <= 1e-8
    assert sc["div_l2"] < 1e-6

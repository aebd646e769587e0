"""Pins of the oracle's immersed-boundary operators and functionals (no GPU).

Every expected value here is computed in numpy from the geometry alone, with
its own circle intersection (the chord x = xc +- sqrt(R^2 - (y - yc)^2) for a
horizontal ray, likewise for a vertical one) instead of the quadratic in the
ray parameter the oracle restates, so that a wrong closure distance, a swapped
neighbour, a wrong wall factor or a wrong row scale fails:

* the u, v, w momentum rows of every active face, cut-cell closures
  c h / d_P included (immersed_operators.hpp:1-175), after the l1 row scaling
  (immersed_operators.hpp:737-748);
* the projection divergence D, cut-cell entries +-alpha / (beta h), and its
  inlet vector (immersed_operators.hpp:179-242);
* the centred convection on arbitrary linear u, v, w fields, where the
  discretisation is exact, and its one-sided fall-back at the walls
  (gpu_dense_ops.hpp:1-14);
* the drag functional: a linear pressure field gives the summation-by-parts
  closed form of the pressure force, a uniform u field gives rho dV times
  the cut-cell closure coefficients of the u rows.
"""
import numpy as np
import pytest

import oracle as O
from helpers import SMALL, csr_1based

GEOMS = {
    "small": SMALL,
    "offset": dict(nx=40, ny=20, nz=6, Lx=1.0, xcirc=0.41, ycirc=0.23, Rcirc=0.061, Um=0.45, dt=2e-3),
}


def P(kw):
    return O.default_params(**kw)


def full(kw):
    d = {k: getattr(O.default_params(), k) for k, _ in O.PARAM_FIELDS}
    d.update(kw)
    return d


@pytest.fixture(scope="module", params=sorted(GEOMS))
def geom(request):
    kw = GEOMS[request.param]
    return O.Setup(P(kw)), full(kw)


def chord_distance(xP, yP, axis, sign, h, xc, yc, R):
    """Distance from (xP, yP) along +-x (axis 0) or +-y (axis 1) to the first
    crossing of the circle within h, or inf."""
    along, across = (xP, yP - yc) if axis == 0 else (yP, xP - xc)
    centre = xc if axis == 0 else yc
    s2 = R * R - across * across
    if s2 < 0.0:
        return np.inf
    s = np.sqrt(s2)
    best = np.inf
    for hit in (centre - s, centre + s):
        d = sign * (hit - along)
        if 0.0 <= d <= h:
            best = min(best, d)
    return best


def expected_rows(S, g, comp):
    """Unscaled 7-point rows [diag, W, E, S, N, B, T] and the cut-closure part of
    the diagonal for every face of the solved box of comp, from the rules of
    immersed_operators.hpp:1-175 with numpy circle crossings."""
    nx, ny, nz = g["nx"], g["ny"], g["nz"]
    dx, dy, dz = g["Lx"] / nx, g["Ly"] / ny, g["Lz"] / nz
    cx, cy, cz = g["nu"] / dx ** 2, g["nu"] / dy ** 2, g["nu"] / dz ** 2
    at = g["rho"] / g["dt"]
    xc, yc, R = g["xcirc"], g["ycirc"], g["Rcirc"]
    if comp == "u":
        act = S.get("active_u").reshape(nz, ny, nx + 1).astype(bool)
        box = (range(nz), range(ny), range(1, nx))
        pos = lambda i, j: (i * dx, (j + 0.5) * dy)
        xlo, xhi = 1, nx - 1
        ylo, yhi = 0, ny - 1
        zlo, zhi = 0, nz - 1
        # boundary factors: (west, east, south, north, bottom, top); None = no term
        wall = dict(W=1.0, E=None, S=2.0, N=2.0, B=2.0, T=2.0)
    elif comp == "v":
        act = S.get("active_v").reshape(nz, ny + 1, nx).astype(bool)
        box = (range(nz), range(1, ny), range(nx))
        pos = lambda i, j: ((i + 0.5) * dx, j * dy)
        xlo, xhi = 0, nx - 1
        ylo, yhi = 1, ny - 1
        zlo, zhi = 0, nz - 1
        wall = dict(W=2.0, E=None, S=1.0, N=1.0, B=2.0, T=2.0)
    else:
        act = S.get("active_w").reshape(nz + 1, ny, nx).astype(bool)
        box = (range(1, nz), range(ny), range(nx))
        pos = lambda i, j: ((i + 0.5) * dx, (j + 0.5) * dy)
        xlo, xhi = 0, nx - 1
        ylo, yhi = 0, ny - 1
        zlo, zhi = 1, nz - 1
        wall = dict(W=2.0, E=None, S=2.0, N=2.0, B=1.0, T=1.0)
    shape = tuple(len(r) for r in box)
    rows = np.zeros(shape + (7,))
    cut = np.zeros(shape)
    for a, k in enumerate(box[0]):
        for b, j in enumerate(box[1]):
            for c, i in enumerate(box[2]):
                if not act[k, j, i]:
                    continue
                xP, yP = pos(i, j)
                row = np.zeros(7)
                diag, cutP = at, 0.0
                nbrs = (
                    (1, "W", i > xlo and act[k, j, i - 1], i == xlo, cx, dx, 0, -1),
                    (2, "E", i < xhi and act[k, j, i + 1], i == xhi, cx, dx, 0, +1),
                    (3, "S", j > ylo and act[k, j - 1, i], j == ylo, cy, dy, 1, -1),
                    (4, "N", j < yhi and act[k, j + 1, i], j == yhi, cy, dy, 1, +1),
                    (5, "B", k > zlo and act[k - 1, j, i], True, cz, dz, 2, -1),
                    (6, "T", k < zhi and act[k + 1, j, i], True, cz, dz, 2, +1),
                )
                for slot, name, open_, boundary, coef, h, axis, sign in nbrs:
                    if open_:
                        row[slot] = -coef
                        diag += coef
                    elif boundary:
                        if wall[name] is not None:
                            diag += coef * wall[name]
                    else:  # a neighbour blocked by the cylinder (x or y only: the cylinder is full height)
                        d = chord_distance(xP, yP, axis, sign, h, xc, yc, R)
                        term = coef * h / d
                        diag += term
                        cutP += term
                row[0] = diag
                rows[a, b, c] = row
                cut[a, b, c] = cutP
    return rows, cut, act[tuple(slice(r.start, r.stop) for r in box)]


def box_ids(S, g, comp):
    nx, ny, nz = g["nx"], g["ny"], g["nz"]
    if comp == "u":
        return S.get("id_u").reshape(nz, ny, nx + 1)[:, :, 1:nx]
    if comp == "v":
        return S.get("id_v").reshape(nz, ny + 1, nx)[:, 1:ny, :]
    return S.get("id_w").reshape(nz + 1, ny, nx)[1:nz]


@pytest.mark.parametrize("comp", ["u", "v", "w"])
def test_momentum_rows_with_cylinder_closures(geom, comp):
    S, g = geom
    rows, cut, act = expected_rows(S, g, comp)
    got = S.get("A%s_vals" % comp).reshape(rows.shape)
    want = rows / np.abs(rows).sum(axis=-1, keepdims=True).clip(min=1e-300)
    # every active row: the l1-scaled closed form
    err = np.abs(got[act] - want[act]).max()
    assert err < 1e-14, err
    # the rows with cut closures are there and are not wall rows
    assert (cut[act] > 0).sum() >= 8
    # the 1-based scale vector holds 1 / sum |a| in id order (the two circle
    # crossings round differently: 1e-12 covers a closure distance of 1e-3 h)
    # (w's wall faces carry ids but no row: their scale is 1)
    scale = S.get("scale_" + comp)
    ids = box_ids(S, g, comp)[act]
    assert np.allclose(scale[ids], 1.0 / np.abs(rows[act]).sum(axis=-1), rtol=1e-12, atol=0)
    rest = np.setdiff1d(np.arange(1, len(scale)), ids)
    assert np.all(scale[rest] == 1.0) and (comp == "w") == (len(rest) > 0)


def test_a_plausible_closure_mistake_is_caught(geom):
    """The same comparison, against rows built with the closure distance taken
    to the neighbour face centre (d = h) instead of the circle, must fail."""
    S, g = geom
    rows, cut, act = expected_rows(S, g, "u")
    wrong = rows.copy()
    h_terms = np.zeros_like(cut)
    # replace every cut term c h / d by c (distance h)
    nx, ny = g["nx"], g["ny"]
    cx = g["nu"] / (g["Lx"] / nx) ** 2
    has_cut = cut > 0
    h_terms[has_cut] = cut[has_cut]
    wrong[..., 0] -= h_terms
    wrong[..., 0] += np.where(has_cut, cx, 0.0)
    got = S.get("Au_vals").reshape(rows.shape)
    want = wrong / np.abs(wrong).sum(axis=-1, keepdims=True).clip(min=1e-300)
    assert np.abs(got[act] - want[act]).max() > 1e-6


def test_divergence_cut_cells_and_inlet_vector(geom):
    S, g = geom
    nx, ny, nz = g["nx"], g["ny"], g["nz"]
    dx, dy, dz = g["Lx"] / nx, g["Ly"] / ny, g["Lz"] / nz
    H, Um = g["Ly"], g["Um"]
    beta = S.get("beta2d").reshape(ny, nx)
    au = S.get("alpha_u2d").reshape(ny, nx + 1)
    av = S.get("alpha_v2d").reshape(ny + 1, nx)
    fl = S.get("fluid_p").reshape(nz, ny, nx).astype(bool)
    idp = S.get("id_p").reshape(nz, ny, nx)
    iqu = S.get("id_qu").reshape(nz, ny, nx + 1)
    iqv = S.get("id_qv").reshape(nz, ny + 1, nx)
    iqw = S.get("id_qw").reshape(nz + 1, ny, nx)
    nQu, nQv, nQw = S.count("Nqu"), S.count("Nqv"), S.count("Nqw")
    Np = S.count("Np")
    want = np.zeros((Np, nQu + nQv + nQw))
    dbc = np.zeros(Np)
    n_cut = 0
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                if not fl[k, j, i]:
                    continue
                r = idp[k, j, i] - 1
                b = beta[j, i]
                n_cut += b < 1.0
                faces = [(iqu[k, j, i + 1], 0, au[j, i + 1] / (b * dx))]
                if i > 0:
                    faces.append((iqu[k, j, i], 0, -au[j, i] / (b * dx)))
                else:
                    y, z = (j + 0.5) * dy, (k + 0.5) * dz
                    dbc[r] = -(16.0 * Um * y * z * (H - y) * (H - z) / H ** 4) / (b * dx)
                if j < ny - 1:
                    faces.append((iqv[k, j + 1, i], nQu, av[j + 1, i] / (b * dy)))
                if j > 0:
                    faces.append((iqv[k, j, i], nQu, -av[j, i] / (b * dy)))
                if k < nz - 1:
                    faces.append((iqw[k + 1, j, i], nQu + nQv, 1.0 / (b * dz)))
                if k > 0:
                    faces.append((iqw[k, j, i], nQu + nQv, -1.0 / (b * dz)))
                for fid, off, val in faces:
                    if fid > 0:  # a face without a projection unknown carries no entry
                        want[r, off + fid - 1] += val
    assert n_cut > 0
    D = csr_1based(S.get("D_rowptr"), S.get("D_colind"), S.get("D_vals"), nQu + nQv + nQw).toarray()
    # every entry is one division, evaluated in the same order: exact
    assert np.array_equal(D, want)
    assert np.array_equal(S.get("div_bc")[1:], dbc)


def _open_channel(nx=20, ny=10, nz=8):
    return dict(nx=nx, ny=ny, nz=nz, Lx=1.0, xcirc=5.0, ycirc=5.0, Rcirc=0.01)


def _centres(g):
    nx, ny, nz = g["nx"], g["ny"], g["nz"]
    dx, dy, dz = g["Lx"] / nx, g["Ly"] / ny, g["Lz"] / nz
    xf, yf, zf = np.arange(nx + 1) * dx, np.arange(ny + 1) * dy, np.arange(nz + 1) * dz
    xc, yc, zc = (np.arange(nx) + 0.5) * dx, (np.arange(ny) + 0.5) * dy, (np.arange(nz) + 0.5) * dz
    return dict(u=(xf, yc, zc), v=(xc, yf, zc), w=(xc, yc, zf))


def _lin(coef, xyz):
    x, y, z = xyz
    return coef[0] + coef[1] * x[None, None, :] + coef[2] * y[None, :, None] + coef[3] * z[:, None, None]


LIN = dict(u=(0.31, 0.7, -0.45, 0.2), v=(-0.12, 0.25, 0.6, -0.33), w=(0.05, -0.4, 0.15, 0.55))


def test_convection_exact_on_linear_fields_all_components():
    kw = _open_channel()
    g = full(kw)
    s = O.Sim(P(kw))
    pos = _centres(g)
    fields = {c: _lin(LIN[c], pos[c]) for c in "uvw"}
    for c in "uvw":
        s.set(c, fields[c].reshape(-1))
    s.phase(0)
    s.phase(1)  # bc cleanup rewrites the boundary faces, then convection
    nx, ny, nz = g["nx"], g["ny"], g["nz"]
    grad = {c: np.array(LIN[c][1:]) for c in "uvw"}
    # velocities at each component's faces: the linear fields evaluated there
    for c, sl in (("u", np.s_[1:nz - 1, 1:ny - 1, 2:nx - 1]),
                  ("v", np.s_[1:nz - 1, 2:ny - 1, 1:nx - 1]),
                  ("w", np.s_[2:nz - 1, 1:ny - 1, 1:nx - 1])):
        vel = [_lin(LIN[q], pos[c]) for q in "uvw"]
        want = sum(vel[d] * grad[c][d] for d in range(3))
        got = s.get("conv_" + c).reshape(want.shape)
        scale = np.abs(want[sl]).max()
        assert np.abs(got[sl] - want[sl]).max() < 1e-13 * scale, c
        # a transposed operand (the x-advecting speed on the y-derivative) must fail
        wrong = vel[1] * grad[c][0] + vel[0] * grad[c][1] + vel[2] * grad[c][2]
        assert np.abs(got[sl] - wrong[sl]).max() > 1e-3 * scale


def test_convection_wall_fallback_is_one_sided_half():
    """At a wall the missing neighbour takes the centre value (gpu_dense_ops.hpp:1-3),
    so the centred difference becomes half the one-sided one."""
    kw = _open_channel()
    g = full(kw)
    nx, ny, nz = g["nx"], g["ny"], g["nz"]
    dy = g["Ly"] / ny
    pos = _centres(g)
    s = O.Sim(P(kw))
    a2 = 0.8
    u = _lin((0.0, 0.0, a2, 0.0), pos["u"])  # u = a2 y
    v = np.full((nz, ny + 1, nx), 0.3)      # uniform v on the interior faces, zeroed at the walls by the cleanup
    s.set("u", u.reshape(-1))
    s.set("v", v.reshape(-1))
    s.set("w", np.zeros(nx * ny * (nz + 1)))
    s.phase(0)
    s.phase(1)
    cu = s.get("conv_u").reshape(nz, ny, nx + 1)
    # south wall row j = 0: vadv = (0 + 0.3 + 0 + 0.3) / 4, du/dy = a2 dy / (2 dy)
    assert np.allclose(cu[:, 0, 2:nx - 1], 0.15 * a2 / 2, rtol=1e-13)
    assert np.allclose(cu[:, ny - 1, 2:nx - 1], 0.15 * a2 / 2, rtol=1e-13)
    assert np.allclose(cu[:, 1:ny - 1, 2:nx - 1], 0.3 * a2, rtol=1e-13)
    del dy


def test_convection_zero_on_inactive_faces(geom):
    S, g = geom
    kw = {k: g[k] for k in ("nx", "ny", "nz", "Lx", "xcirc", "ycirc", "Rcirc", "Um", "dt")}
    s = O.Sim(P(dict(kw, init_profile=1, noise_amp=0.01, seed=3)))
    s.phase(0)
    s.phase(1)
    for c, m in (("u", "active_u"), ("v", "active_v")):
        conv = s.get("conv_" + c)
        act = S.get(m).astype(bool)
        assert np.all(conv[~act] == 0.0) and np.any(conv[act] != 0.0)


def _force(kw, u=None, p=None):
    """fx, fy of the oracle for the given fields with every other field zero."""
    g = full(kw)
    nx, ny, nz = g["nx"], g["ny"], g["nz"]
    s = O.Sim(P(kw))
    s.set("u", np.zeros((nx + 1) * ny * nz) if u is None else u.reshape(-1))
    s.set("v", np.zeros(nx * (ny + 1) * nz))
    s.set("w", np.zeros(nx * ny * (nz + 1)))
    s.set("p", np.zeros(nx * ny * nz) if p is None else p.reshape(-1))
    s.phase(7)
    sc = s.scalars()
    return sc["fx"], sc["fy"]


def _by_faces(alpha, fl, pc, slope_h, axis):
    """Summation by parts of sum_c p_c (alpha_lo(c) - alpha_hi(c)) over the fluid
    cells c of each line along axis: sum over the faces f of alpha_f (P_hi - P_lo),
    P = p on fluid cells and 0 elsewhere.  Between two fluid cells the jump of the
    linear field is slope_h exactly; at a face with one fluid side it is +-p of
    that side; the first and last faces see +p_first and -p_last."""
    if axis == 1:
        alpha, fl, pc = alpha.T, fl.T, pc.T
    inner = alpha[:, 1:-1]
    both = fl[:, :-1] & fl[:, 1:]
    hi_only = ~fl[:, :-1] & fl[:, 1:]
    lo_only = fl[:, :-1] & ~fl[:, 1:]
    total = slope_h * inner[both].sum()
    total += (inner[hi_only] * pc[:, 1:][hi_only]).sum() - (inner[lo_only] * pc[:, :-1][lo_only]).sum()
    total += (alpha[:, 0] * pc[:, 0] * fl[:, 0]).sum() - (alpha[:, -1] * pc[:, -1] * fl[:, -1]).sum()
    return total


def test_drag_of_linear_pressure_is_summation_by_parts(geom):
    """F_x = sum_c p_c (alpha_west - alpha_east) dy dz and F_y likewise over the
    fluid cells; for p = a + b x + c y the face-wise sum (summation by parts)
    gives a closed form in the face fractions alone."""
    S, g = geom
    kw = {k: g[k] for k in ("nx", "ny", "nz", "Lx", "xcirc", "ycirc", "Rcirc", "Um", "dt")}
    nx, ny, nz = g["nx"], g["ny"], g["nz"]
    dx, dy, dz = g["Lx"] / nx, g["Ly"] / ny, g["Lz"] / nz
    fl2 = S.get("fluid2d").reshape(ny, nx).astype(bool)
    au = S.get("alpha_u2d").reshape(ny, nx + 1)
    av = S.get("alpha_v2d").reshape(ny + 1, nx)
    a0, b, c = 0.7, -1.3, 2.1
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dy
    pc = a0 + b * x[None, :] + c * y[:, None]
    p = np.ascontiguousarray(np.broadcast_to(pc * fl2, (nz, ny, nx)))
    fx, fy = _force(kw, p=p)
    want_x = dy * dz * nz * _by_faces(au, fl2, pc, b * dx, 0)
    want_y = dx * dz * nz * _by_faces(av, fl2, pc, c * dy, 1)
    assert abs(fx - want_x) < 1e-12 * abs(want_x) and want_x != 0.0
    assert abs(fy - want_y) < 1e-12 * abs(want_y) and want_y != 0.0
    # swapped face areas (dx dz on the x force) would be caught
    assert abs(fx - want_x * dx / dy) > 1e-3 * abs(want_x)
    # minus the gradient times the blocked volume, to the accuracy of a circle
    # 2.4 to 3.4 cells in radius
    blocked = np.pi * g["Rcirc"] ** 2 * g["Lz"]
    assert abs(fx + b * blocked) < 0.4 * abs(b) * blocked
    assert abs(fy + c * blocked) < 0.4 * abs(c) * blocked
    # a constant pressure exerts only the one-sided faces' force
    fx0, fy0 = _force(kw, p=np.ascontiguousarray(np.broadcast_to(3.0 * fl2, (nz, ny, nx))))
    c3 = np.full_like(pc, 3.0)
    assert abs(fx0 - dy * dz * nz * _by_faces(au, fl2, c3, 0.0, 0)) < 1e-13
    assert abs(fy0 - dx * dz * nz * _by_faces(av, fl2, c3, 0.0, 1)) < 1e-13


def test_viscous_drag_weights_are_the_u_closures(geom):
    """For u = U on every active face (v = w = p = 0) the viscous x force is
    rho dV U times the sum of the u rows' cut-cell closure coefficients c h / d."""
    S, g = geom
    kw = {k: g[k] for k in ("nx", "ny", "nz", "Lx", "xcirc", "ycirc", "Rcirc", "Um", "dt")}
    nx, ny, nz = g["nx"], g["ny"], g["nz"]
    dV = g["Lx"] / nx * g["Ly"] / ny * g["Lz"] / nz
    _, cut, act = expected_rows(S, g, "u")
    U = 0.37
    u = np.zeros((nz, ny, nx + 1))
    actu = S.get("active_u").reshape(nz, ny, nx + 1).astype(bool)
    u[actu] = U
    fx, fy = _force(kw, u=u)
    want = g["rho"] * dV * U * cut[act].sum()
    assert abs(fx - want) < 1e-12 * abs(want) and want > 0.0
    assert fy == 0.0

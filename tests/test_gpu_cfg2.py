"""Whole-step parity at the production grid 512x96x96 (Um 2.25, Re ~100, dt 2.5e-4,
GMRES(30), inlet profile + seeded noise), on the kernels the production run uses:
z-marching tiles on the hot copies of the four operators (the pressure operator's
with its flagged rows), the tiled V-cycle levels and the one-point coarse levels.

The oracle's state after three steps at rtol 1e-10 is committed as a fixture
(tests/golden/cfg2_3steps.npz, written by tests/golden/make_cfg2_fixture.py):
two planes per field, the L2 norm of every z-plane and of the whole field, c_D
and c_L.  Bars: 1e-9 relative L2 on planes and norms, 1e-7 relative on c_D and
c_L (c_L against c_D's magnitude, as it is near zero).  A second run must give
the same fields bit for bit."""
import hashlib
import os

import numpy as np
import pytest

import fpc_amd as F
from cfg2_case import CFG2_PARITY, NSTEPS, field_views
from helpers import rel_l2

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cfg2_3steps.npz")
FIELDS = ("u", "v", "w", "p")


def _run():
    p = CFG2_PARITY
    g = F.Sim(F.default_params(**p))
    its = []
    for _ in range(NSTEPS):
        g.step(1)
        its.append([int(g.scalar("its_" + c)) for c in "uvwp"])
    views, digests = {}, {}
    for f in FIELDS:
        a = g.get(f)
        views.update(field_views(f, a, p["nx"], p["ny"], p["nz"]))
        digests[f] = hashlib.sha256(a.tobytes()).hexdigest()
    scal = {k: g.scalar(k) for k in ("cd", "cl", "relres_p")}
    ext = {c: (g.scalar("zext_m_" + c), g.scalar("zext_flagged_" + c)) for c in "uvwp"}
    g.close()
    return views, digests, scal, np.array(its), ext


@pytest.fixture(scope="module")
def runs(torch_cuda):
    assert os.path.exists(FIXTURE), FIXTURE
    first = _run()
    second = _run()
    return first, second


def test_fields_match_oracle_fixture(runs):
    (views, _, scal, its, ext), _ = runs
    ref = np.load(FIXTURE)
    # the production kernels ran: every operator through its hot copy, flagged pressure rows
    assert all(m >= 1 for m, _ in ext.values()), ext
    assert ext["p"][1] > 0, ext
    assert np.array_equal(its, ref["its"]), (its, ref["its"])
    for f in FIELDS:
        for view in ("_zmid", "_xcyl", "_knorm"):
            err = rel_l2(views[f + view], ref[f + view])
            assert err < 1e-9, (f + view, err)
        assert abs(views[f + "_norm"] - ref[f + "_norm"]) <= 1e-9 * ref[f + "_norm"], f
    cd, cl = float(ref["cd"]), float(ref["cl"])
    assert abs(scal["cd"] - cd) <= 1e-7 * abs(cd), (scal["cd"], cd)
    assert abs(scal["cl"] - cl) <= 1e-7 * abs(cd), (scal["cl"], cl)
    assert scal["relres_p"] <= CFG2_PARITY["p_tol"]


def test_second_run_bitwise(runs):
    (_, d1, s1, i1, _), (_, d2, s2, i2, _) = runs
    assert d1 == d2
    assert np.array_equal(i1, i2)
    assert s1 == s2

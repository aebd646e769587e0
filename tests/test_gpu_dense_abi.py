"""Element-wise parity of every function of include/fpc_dense.h against the oracle.

Layer 1 (one function per reference kernel of gpu_dense_ops.hpp:5-106) is called
in the reference's launch order of one time step (main.cu:256-350) on device
copies of the oracle's state, each call fed the oracle's own inputs, so every
output is compared on identical inputs:

* copies, masks, packs, unpacks and scatters bit for bit;
* convection, the CSR SpMV and the rhs builders within 1e-14 of the largest
  value (the device contracts a*b + c into fused multiply-adds, as nvcc does
  for the reference kernels; the oracle rounds each operation).

Layer 2 (the fused stages) is called on the operator data of a simulation
(fpc_sim_flow_ops) and compared with the oracle's phases in the same way.
"""
import ctypes

import numpy as np
import pytest

import fpc_amd as F
import oracle as O
from helpers import SMALL

pytestmark = pytest.mark.gpu

KW = dict(SMALL, init_profile=1, noise_amp=1e-3 * 0.45, seed=21, inlet_ramp_steps=4, vel_tol=1e-12, p_tol=1e-12)


def _close(a, b, what, rel=1e-14):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, what
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b).max() / scale
    assert err <= rel, (what, err)


def _bits(a, b, what):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.shape == b.shape and a.dtype == b.dtype, what
    if a.dtype == np.float64:
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), what
    else:
        assert np.array_equal(a, b), what


@pytest.fixture(scope="module")
def state(torch_cuda):
    """The oracle's arrays around the phases of its second step (the first step makes
    the state non-trivial: v, w and p non-zero)."""
    p = O.default_params(**KW)
    S = O.Setup(p)
    o = O.Sim(p)
    o.step(1)
    st = {"setup": S, "params": p}
    st["pre"] = {f: o.get(f) for f in "uvwp"}
    o.phase(0)
    st["uin_step"], st["ramp"] = o.get("uin_step"), o.scalar("ramp")
    o.phase(1)
    for n in ("u", "v", "w", "conv_u", "conv_v", "conv_w", "gpx", "gpy", "gpz", "rhs_u", "rhs_v", "rhs_w", "x_u",
              "x_v", "x_w"):
        st["ph1_" + n] = o.get(n)
    o.phase(2)
    for n in ("x_u", "x_v", "x_w"):
        st["ph2_" + n] = o.get(n)
    o.phase(3)
    for n in ("u_tilde", "v_tilde", "w_tilde"):
        st["ph3_" + n] = o.get(n)
    o.phase(4)
    for n in ("qt", "Dqt", "rhs_p", "x_p"):
        st["ph4_" + n] = o.get(n)
    o.phase(5)
    st["ph5_x_p"] = o.get("x_p")
    o.phase(6)
    for n in ("u", "v", "w", "p", "qt"):
        st["ph6_" + n] = o.get(n)
    o.phase(7)
    st["ph7"] = o.scalars()
    return st


class Dev:
    """Device copies (torch CUDA tensors) and raw pointers for the C ABI."""

    def __init__(self):
        import torch
        self.torch = torch
        self.t = {}

    def put(self, name, arr):
        a = np.ascontiguousarray(arr)
        self.t[name] = self.torch.from_numpy(a.copy()).to("cuda")
        return self.p(name)

    def empty(self, name, n, fill=np.nan):
        self.t[name] = self.torch.full((int(n),), fill, dtype=self.torch.float64, device="cuda")
        return self.p(name)

    def p(self, name):
        return ctypes.c_void_p(self.t[name].data_ptr())

    def get(self, name):
        self.torch.cuda.synchronize()
        return self.t[name].cpu().numpy()

    def sync(self):
        self.torch.cuda.synchronize()


def _call(fn, *args):
    import torch
    torch.cuda.synchronize()
    F.check(fn(*args), fn.__name__)
    F.check(F.lib().fpc_DeviceSynchronize(), "sync")


def test_layer1_in_reference_launch_order(state):
    L = F.lib()
    S, par = state["setup"], state["params"]
    nx, ny, nz = par.nx, par.ny, par.nz
    dx, dy, dz = par.Lx / nx, par.Ly / ny, par.Lz / nz
    N = nx * ny * nz
    Nu, Nv, Nw = (nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)
    Np, Nqu, Nqv, Nqw = S.count("Np"), S.count("Nqu"), S.count("Nqv"), S.count("Nqw")
    Nq = Nqu + Nqv + Nqw
    d = Dev()
    for n in ("active_u", "active_v", "active_w", "fluid_p", "id_u", "id_v", "id_w", "id_p", "id_qu", "id_qv",
              "id_qw", "scale_u", "scale_v", "scale_w", "rhs_u_bc", "rhs_v_bc", "rhs_w_bc", "div_bc", "D_rowptr",
              "D_colind", "D_vals", "G_rowptr", "G_colind", "G_vals"):
        d.put(n, S.get(n))
    pre = state["pre"]
    for f in "uvwp":
        d.put(f, pre[f])
    d.put("uin_step", state["uin_step"])

    # main.cu:256 BC cleanup, then :257-259 convection
    _call(L.fpc_apply_bc_cleanup, d.p("u"), d.p("v"), d.p("w"), d.p("active_u"), d.p("active_v"), d.p("active_w"),
          d.p("uin_step"), nx, ny, nz)
    for f in "uvw":
        _bits(d.get(f), state["ph1_" + f], "bc_cleanup " + f)
    for f, n in (("u", Nu), ("v", Nv), ("w", Nw)):
        d.empty("conv_" + f, n)
        _call(getattr(L, "fpc_conv_" + f), d.p("conv_" + f), d.p("u"), d.p("v"), d.p("w"), d.p("active_" + f), nx, ny,
              nz, dx, dy, dz)
        _close(d.get("conv_" + f), state["ph1_conv_" + f], "conv_" + f)

    # main.cu:262-271 memsets, pack p, G p (CSR), unpack onto the faces
    for n, ln in (("pvec", Np + 1), ("gp", Nq + 1), ("gpx", Nu), ("gpy", Nv), ("gpz", Nw)):
        d.empty(n, ln)
        _call(L.fpc_zero_doubles, d.p(n), ln)
    _call(L.fpc_pack_pressure_active, d.p("pvec"), d.p("p"), d.p("fluid_p"), d.p("id_p"), N)
    pvec = np.zeros(Np + 1)
    fl, idp = S.get("fluid_p").astype(bool), S.get("id_p")
    pvec[idp[fl & (idp > 0)]] = pre["p"][fl & (idp > 0)]
    _bits(d.get("pvec"), pvec, "pack_pressure_active")
    _call(L.fpc_spmv_csr_1based, Nq, d.p("G_rowptr"), d.p("G_colind"), d.p("G_vals"), d.p("pvec"), d.p("gp"))
    gp_ref = O.csr_spmv_1based(S.get("G_rowptr"), S.get("G_colind"), S.get("G_vals"), pvec)
    _close(d.get("gp"), gp_ref, "spmv G p")
    _call(L.fpc_unpack_gp_u, d.p("gp"), d.p("gpx"), d.p("id_qu"), nx, ny, nz)
    _call(L.fpc_unpack_gp_v, d.p("gp"), d.p("gpy"), d.p("id_qv"), nx, ny, nz, Nqu)
    _call(L.fpc_unpack_gp_w, d.p("gp"), d.p("gpz"), d.p("id_qw"), nx, ny, nz, Nqu, Nqv)
    for f, g in (("gpx", "u"), ("gpy", "v"), ("gpz", "w")):
        _close(d.get(f), state["ph1_" + f], "unpack " + f)

    # main.cu:272-274 momentum rhs / x0 boxes, from the oracle's own conv and G p
    boxes = {"u": (nx - 1) * ny * nz, "v": nx * (ny - 1) * nz, "w": nx * ny * (nz - 1)}
    for f, g in (("u", "gpx"), ("v", "gpy"), ("w", "gpz")):
        d.put("oconv_" + f, state["ph1_conv_" + f])
        d.put("o" + g, state["ph1_" + g])
        d.empty("rhs_" + f, boxes[f])
        d.empty("x_" + f, boxes[f])
        _call(getattr(L, "fpc_build_%s_full" % f), d.p("rhs_" + f), d.p("x_" + f), d.p(f), d.p("oconv_" + f),
              d.p("o" + g), d.p("active_" + f), d.p("id_" + f), d.p("rhs_%s_bc" % f), d.p("scale_" + f), nx, ny, nz,
              par.rho, par.dt, 1)
        _close(d.get("rhs_" + f), state["ph1_rhs_" + f], "build rhs " + f)
        _bits(d.get("x_" + f), state["ph1_x_" + f], "build x0 " + f)

    # main.cu:307-312 scatter the solutions, BC cleanup of the tilde fields
    for f, n in (("u", Nu), ("v", Nv), ("w", Nw)):
        d.put("sol_" + f, state["ph2_x_" + f])
        d.empty(f + "t", n)
        _call(getattr(L, "fpc_scatter_" + f), d.p(f + "t"), d.p("sol_" + f), d.p(f), nx, ny, nz)
    _call(L.fpc_apply_bc_cleanup, d.p("ut"), d.p("vt"), d.p("wt"), d.p("active_u"), d.p("active_v"),
          d.p("active_w"), d.p("uin_step"), nx, ny, nz)
    for f in "uvw":
        _bits(d.get(f + "t"), state["ph3_%s_tilde" % f], "scatter " + f)

    # main.cu:314-321 pack q~, D q~ (CSR), pressure rhs
    d.empty("q", Nq + 1)
    _call(L.fpc_zero_doubles, d.p("q"), Nq + 1)
    _call(L.fpc_pack_projection_q_u, d.p("q"), d.p("ut"), d.p("id_qu"), nx, ny, nz)
    _call(L.fpc_pack_projection_q_v, d.p("q"), d.p("vt"), d.p("id_qv"), nx, ny, nz, Nqu)
    _call(L.fpc_pack_projection_q_w, d.p("q"), d.p("wt"), d.p("id_qw"), nx, ny, nz, Nqu, Nqv)
    _bits(d.get("q"), state["ph4_qt"], "pack_projection_q")
    d.empty("Dqt", Np + 1)
    _call(L.fpc_zero_doubles, d.p("Dqt"), Np + 1)
    _call(L.fpc_spmv_csr_1based, Np, d.p("D_rowptr"), d.p("D_colind"), d.p("D_vals"), d.p("q"), d.p("Dqt"))
    _close(d.get("Dqt"), state["ph4_Dqt"], "spmv D q")
    d.put("oDqt", state["ph4_Dqt"])
    d.empty("rhs_p", N)
    d.empty("x_p", N)
    _call(L.fpc_build_pressure_rhs_full, d.p("rhs_p"), d.p("x_p"), d.p("oDqt"), d.p("div_bc"), d.p("fluid_p"),
          d.p("id_p"), N, par.rho, par.dt, state["ramp"])
    _close(d.get("rhs_p"), state["ph4_rhs_p"], "build pressure rhs")
    _bits(d.get("x_p"), state["ph4_x_p"], "pressure x0")

    # main.cu:338-350 pack phi, G phi, q correction, u = u~, unpack q, BC cleanup, p += phi
    d.put("xp_sol", state["ph5_x_p"])
    d.empty("phi", Np + 1)
    _call(L.fpc_zero_doubles, d.p("phi"), Np + 1)
    _call(L.fpc_pack_phi_active, d.p("phi"), d.p("xp_sol"), d.p("fluid_p"), d.p("id_p"), N)
    phi = np.zeros(Np + 1)
    phi[idp[fl & (idp > 0)]] = state["ph5_x_p"][fl & (idp > 0)]
    _bits(d.get("phi"), phi, "pack_phi_active")
    d.empty("gphi", Nq + 1)
    _call(L.fpc_zero_doubles, d.p("gphi"), Nq + 1)
    _call(L.fpc_spmv_csr_1based, Nq, d.p("G_rowptr"), d.p("G_colind"), d.p("G_vals"), d.p("phi"), d.p("gphi"))
    gphi = O.csr_spmv_1based(S.get("G_rowptr"), S.get("G_colind"), S.get("G_vals"), phi)
    _close(d.get("gphi"), gphi, "spmv G phi")
    d.put("ogphi", gphi)
    d.put("q6", state["ph4_qt"])
    _call(L.fpc_apply_q_correction, d.p("q6"), d.p("ogphi"), par.dt / par.rho, Nq + 1)
    _close(d.get("q6"), state["ph6_qt"], "q correction")
    d.put("oq", state["ph6_qt"])
    for f in "uvw":
        d.put(f + "n", state["ph3_%s_tilde" % f])
    _call(L.fpc_unpack_projection_q_u, d.p("oq"), d.p("un"), d.p("id_qu"), nx, ny, nz)
    _call(L.fpc_unpack_projection_q_v, d.p("oq"), d.p("vn"), d.p("id_qv"), nx, ny, nz, Nqu)
    _call(L.fpc_unpack_projection_q_w, d.p("oq"), d.p("wn"), d.p("id_qw"), nx, ny, nz, Nqu, Nqv)
    _call(L.fpc_apply_bc_cleanup, d.p("un"), d.p("vn"), d.p("wn"), d.p("active_u"), d.p("active_v"),
          d.p("active_w"), d.p("uin_step"), nx, ny, nz)
    for f in "uvw":
        _bits(d.get(f + "n"), state["ph6_" + f], "projected " + f)
    d.put("ophi", phi)
    _call(L.fpc_add_pressure_phi, d.p("p"), d.p("ophi"), d.p("fluid_p"), d.p("id_p"), N)
    _bits(d.get("p"), state["ph6_p"], "add_pressure_phi")


def test_fused_stages_match_oracle_phases(state):
    """fpc_stage_{explicit,scatter,divergence,projection,diagnostics} on a simulation's
    operator data, fed the oracle's state before each phase."""
    L = F.lib()
    par = state["params"]
    nx, ny, nz = par.nx, par.ny, par.nz
    sim = F.Sim(F.default_params(**KW))
    ops = sim.flow_ops()
    F.check(L.fpc_DeviceSynchronize(), "sync")
    d = Dev()
    pre = state["pre"]
    for f in "uvwp":
        d.put(f, pre[f])
    d.put("uin_step", state["uin_step"])
    boxes = {"u": (nx - 1) * ny * nz, "v": nx * (ny - 1) * nz, "w": nx * ny * (nz - 1)}
    for f in "uvw":
        d.empty("rhs_" + f, boxes[f])
        d.empty("x_" + f, boxes[f])
    _call(L.fpc_stage_explicit, ctypes.byref(ops), d.p("u"), d.p("v"), d.p("w"), d.p("p"), d.p("uin_step"),
          d.p("rhs_u"), d.p("x_u"), d.p("rhs_v"), d.p("x_v"), d.p("rhs_w"), d.p("x_w"), 1)
    for f in "uvw":
        _bits(d.get(f), state["ph1_" + f], "explicit bc " + f)
        _close(d.get("rhs_" + f), state["ph1_rhs_" + f], "explicit rhs " + f, rel=1e-13)
        _bits(d.get("x_" + f), state["ph1_x_" + f], "explicit x0 " + f)
    for f in "uvw":
        d.put("sol_" + f, state["ph2_x_" + f])
        d.empty(f + "t", len(pre[f]))
    _call(L.fpc_stage_scatter, ctypes.byref(ops), d.p("sol_u"), d.p("sol_v"), d.p("sol_w"), d.p("u"), d.p("v"),
          d.p("w"), d.p("ut"), d.p("vt"), d.p("wt"), d.p("uin_step"))
    for f in "uvw":
        _bits(d.get(f + "t"), state["ph3_%s_tilde" % f], "stage scatter " + f)
    N = nx * ny * nz
    d.empty("rhs_p", N)
    d.empty("x_p", N)
    _call(L.fpc_stage_divergence, ctypes.byref(ops), d.p("ut"), d.p("vt"), d.p("wt"), state["ramp"], d.p("rhs_p"),
          d.p("x_p"))
    _close(d.get("rhs_p"), state["ph4_rhs_p"], "stage divergence", rel=1e-13)
    _bits(d.get("x_p"), state["ph4_x_p"], "stage divergence x0")
    d.put("phi", state["ph5_x_p"])
    d.empty("qout", ny * nz)
    _call(L.fpc_stage_projection, ctypes.byref(ops), d.p("phi"), d.p("ut"), d.p("vt"), d.p("wt"), d.p("u"), d.p("v"),
          d.p("w"), d.p("p"), d.p("uin_step"), d.p("qout"))
    for f in "uvwp":
        _close(d.get(f), state["ph6_" + f], "stage projection " + f, rel=1e-13)
    d.empty("diag", 8)
    n_p = state["setup"].count("Np")
    _call(L.fpc_stage_diagnostics, ctypes.byref(ops), d.p("u"), d.p("v"), d.p("w"), d.p("p"), d.p("qout"),
          state["ramp"], n_p, d.p("diag"))
    diag = d.get("diag")
    o7 = state["ph7"]
    for q, k in ((1, "umax"), (2, "vmax"), (3, "wmax")):
        assert abs(diag[q] - o7[k]) <= 1e-12 * abs(o7[k]), k
    # divL2 is solver-noise sized after the projection (both below the solve tolerance bound)
    assert diag[0] < 1e-6 and o7["div_l2"] < 1e-6
    assert abs(diag[4] - o7["fx"]) <= 1e-9 * abs(o7["fx"])
    assert abs(diag[5] - o7["fy"]) <= 1e-9 * abs(o7["fx"])

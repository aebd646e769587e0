"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, always as the checker.  The oracle restates the reference algorithm
(JD63021/ibm-3d-fpc main.cu:20-376, gpu_dense_ops.hpp, immersed_*.hpp,
solver_helpers.hpp) in plain C++ (sequential unless set_threads(n > 1)); see fpc_oracle.cpp for the line
citations and for which parts are pinned by what.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libfpc_oracle.so")

PARAM_FIELDS = [
    ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
    ("Lx", ctypes.c_double), ("Ly", ctypes.c_double), ("Lz", ctypes.c_double),
    ("rho", ctypes.c_double), ("nu", ctypes.c_double), ("Um", ctypes.c_double), ("dt", ctypes.c_double),
    ("nsteps", ctypes.c_int), ("print_every", ctypes.c_int), ("inlet_ramp_steps", ctypes.c_int),
    ("use_zero_initial_guess", ctypes.c_int), ("nsub_beta", ctypes.c_int),
    ("beta_tol", ctypes.c_double), ("alpha_tol", ctypes.c_double),
    ("xcirc", ctypes.c_double), ("ycirc", ctypes.c_double), ("Rcirc", ctypes.c_double),
    ("vel_maxit", ctypes.c_int), ("vel_kdim", ctypes.c_int), ("vel_tol", ctypes.c_double),
    ("vel_jacobi_iters", ctypes.c_int),
    ("p_maxit", ctypes.c_int), ("p_kdim", ctypes.c_int), ("p_tol", ctypes.c_double),
    ("p_pc_maxit", ctypes.c_int), ("p_pc_relax_type", ctypes.c_int), ("p_pc_rap_type", ctypes.c_int),
    ("init_profile", ctypes.c_int), ("noise_amp", ctypes.c_double), ("seed", ctypes.c_ulonglong),
]


class OrcParams(ctypes.Structure):
    _fields_ = PARAM_FIELDS


_lib = None


def build():
    """Compile the oracle with its Makefile (g++)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        vp, cp, ll = ctypes.c_void_p, ctypes.c_char_p, ctypes.c_longlong
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_default_params.argtypes = [ctypes.POINTER(OrcParams)]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_coarse_size.argtypes = [ctypes.c_int]
        L.orc_setup_create.argtypes = [ctypes.POINTER(OrcParams)]
        L.orc_setup_create.restype = vp
        L.orc_setup_destroy.argtypes = [vp]
        L.orc_setup_len.argtypes = [vp, cp]
        L.orc_setup_len.restype = ll
        L.orc_setup_elsize.argtypes = [vp, cp]
        L.orc_setup_copy.argtypes = [vp, cp, vp, ll]
        L.orc_setup_int.argtypes = [vp, cp]
        L.orc_setup_int.restype = ll
        L.orc_struct_matvec.argtypes = [ctypes.c_int] * 3 + [dp, dp, dp]
        L.orc_csr_spmv_1based.argtypes = [ctypes.c_int, ip, ip, dp, dp, dp]
        L.orc_gmres.argtypes = [ctypes.c_int] * 3 + [dp, dp, dp, ctypes.c_double] + [ctypes.c_int] * 6 + [ip, dp]
        L.orc_precond_apply.argtypes = [ctypes.c_int] * 3 + [dp, dp, dp] + [ctypes.c_int] * 4
        L.orc_mg_levels.argtypes = [ctypes.c_int] * 3 + [dp, ctypes.c_int, ctypes.c_int, ip, ip, dp]
        L.orc_mg_level_matrix.argtypes = [ctypes.c_int] * 3 + [dp, ctypes.c_int, ctypes.c_int, dp]
        L.orc_sim_create.argtypes = [ctypes.POINTER(OrcParams)]
        L.orc_sim_create.restype = vp
        L.orc_sim_destroy.argtypes = [vp]
        L.orc_sim_step.argtypes = [vp, ctypes.c_int]
        L.orc_sim_phase.argtypes = [vp, ctypes.c_int]
        L.orc_sim_len.argtypes = [vp, cp]
        L.orc_sim_len.restype = ll
        L.orc_sim_copy.argtypes = [vp, cp, vp, ll]
        L.orc_sim_set.argtypes = [vp, cp, vp, ll]
        L.orc_sim_scalar.argtypes = [vp, cp]
        L.orc_sim_scalar.restype = ctypes.c_double
        L.orc_write_vtk_cellcentered.argtypes = [cp] + [vp] * 5 + [ctypes.c_int] * 3 + [ctypes.c_double] * 3
        _lib = L
    return _lib


def set_threads(n):
    """Host threads of the oracle's solver and stage loops (1 = the sequential sums)."""
    lib().orc_set_threads(int(n))


def set_coarse_size(n):
    """Largest coarsest MG level (0 = the size-scaled default rule)."""
    lib().orc_set_coarse_size(int(n))


def default_params(**kw):
    p = OrcParams()
    lib().orc_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


_DT = {1: np.uint8, 4: np.int32, 8: np.float64}


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class Setup:
    """One-time setup arrays restated from the reference (masks, ids, boxes, CSR)."""

    INTS = ("Nu", "Nv", "Nw", "Np", "Nqu", "Nqv", "Nqw", "n_pruned")

    def __init__(self, params):
        self.params = params
        self.h = lib().orc_setup_create(ctypes.byref(params))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_setup_destroy(self.h)
            self.h = None

    def get(self, name):
        n = lib().orc_setup_len(self.h, name.encode())
        if n < 0:
            raise KeyError(name)
        es = lib().orc_setup_elsize(self.h, name.encode())
        dt = np.float64 if name.startswith(("beta", "alpha", "rhs", "div", "scale", "uin")) or name.endswith("vals") else _DT[es]
        out = np.empty(n, dtype=dt)
        rc = lib().orc_setup_copy(self.h, name.encode(), out.ctypes.data, out.nbytes)
        if rc:
            raise RuntimeError(f"orc_setup_copy({name}) = {rc}")
        return out

    def count(self, name):
        return int(lib().orc_setup_int(self.h, name.encode()))


def struct_matvec(nx, ny, nz, vals, x):
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(nx * ny * nz)
    lib().orc_struct_matvec(nx, ny, nz, _dp(vals), _dp(x), _dp(y))
    return y


def csr_spmv_1based(rowptr, colind, vals, x):
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    colind = np.ascontiguousarray(colind, dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(len(rowptr), dtype=np.float64)
    lib().orc_csr_spmv_1based(len(rowptr) - 1, _ip(rowptr), _ip(colind), _dp(vals), _dp(x), _dp(y))
    return y


PRECOND = {"none": 0, "jacobi": 1, "pfmg": 2}


def gmres(nx, ny, nz, vals, b, x0=None, tol=1e-8, maxit=250, kdim=30, precond="jacobi", pc_iters=1,
          relax_type=1, rap_type=1):
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64)
    its = ctypes.c_int(0)
    rel = ctypes.c_double(0.0)
    rc = lib().orc_gmres(nx, ny, nz, _dp(vals), _dp(b), _dp(x), tol, maxit, kdim, PRECOND[precond], pc_iters,
                         relax_type, rap_type, ctypes.byref(its), ctypes.byref(rel))
    return x, int(its.value), float(rel.value), int(rc)


def precond_apply(nx, ny, nz, vals, b, precond="pfmg", pc_iters=1, relax_type=1, rap_type=1):
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    lib().orc_precond_apply(nx, ny, nz, _dp(vals), _dp(b), _dp(x), PRECOND[precond], pc_iters, relax_type, rap_type)
    return x


def mg_levels(nx, ny, nz, vals, rap_type=1, max_levels=0):
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    dims = np.zeros(3 * 64, dtype=np.int32)
    cdir = np.zeros(64, dtype=np.int32)
    om = np.zeros(64, dtype=np.float64)
    nl = lib().orc_mg_levels(nx, ny, nz, _dp(vals), rap_type, max_levels, _ip(dims), _ip(cdir), _dp(om))
    return [(tuple(dims[3 * l:3 * l + 3]), int(cdir[l]), float(om[l])) for l in range(nl)]


def mg_level_matrix(nx, ny, nz, vals, level, relax_type=1):
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    dims = mg_levels(nx, ny, nz, vals)[level][0]
    out = np.empty(7 * int(dims[0]) * int(dims[1]) * int(dims[2]))
    rc = lib().orc_mg_level_matrix(nx, ny, nz, _dp(vals), relax_type, level, _dp(out))
    if rc:
        raise IndexError(level)
    return out


class Sim:
    """The oracle time loop (main.cu:249-363 restated)."""

    SCALARS = ("ramp", "div_l2", "umax", "vmax", "wmax", "relres_u", "relres_v", "relres_w", "relres_p",
               "its_u", "its_v", "its_w", "its_p", "err_u", "err_v", "err_w", "err_p", "step", "fx", "fy", "cd", "cl")

    def __init__(self, params):
        self.params = params
        self.h = lib().orc_sim_create(ctypes.byref(params))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sim_destroy(self.h)
            self.h = None

    def step(self, n=1):
        lib().orc_sim_step(self.h, n)

    def phase(self, ph):
        rc = lib().orc_sim_phase(self.h, ph)
        if rc:
            raise RuntimeError(f"phase {ph} -> {rc}")

    def get(self, name):
        n = lib().orc_sim_len(self.h, name.encode())
        if n < 0:
            raise KeyError(name)
        if name.startswith(("id_",)):
            dt = np.int32
        elif name.startswith(("active", "fluid")):
            dt = np.uint8
        elif name.endswith(("rowptr", "colind")):
            dt = np.int32
        else:
            dt = np.float64
        out = np.empty(n, dtype=dt)
        rc = lib().orc_sim_copy(self.h, name.encode(), out.ctypes.data, out.nbytes)
        if rc:
            raise RuntimeError(f"orc_sim_copy({name}) = {rc}")
        return out

    def set(self, name, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        rc = lib().orc_sim_set(self.h, name.encode(), arr.ctypes.data, arr.nbytes)
        if rc:
            raise RuntimeError(f"orc_sim_set({name}) = {rc}")

    def scalar(self, name):
        return float(lib().orc_sim_scalar(self.h, name.encode()))

    def scalars(self):
        return {k: self.scalar(k) for k in self.SCALARS}


def write_vtk_cellcentered(filename, u, v, w, p, fluid_p, nx, ny, nz, Lx, Ly, Lz):
    """The reference's cell-centred VTK writer restated with fprintf (immersed_operators.hpp:1055-1076)."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, w, p)]
    fl = np.ascontiguousarray(fluid_p, dtype=np.uint8)
    rc = lib().orc_write_vtk_cellcentered(os.fsencode(filename), *[a.ctypes.data for a in arrs], fl.ctypes.data,
                                          nx, ny, nz, Lx, Ly, Lz)
    if rc:
        raise OSError(f"cannot write {filename}")


def vtk_filename_with_step(base, step):
    """solver_helpers.hpp:470-482: "_%06d" inserted before the last extension."""
    dot = base.rfind(".")
    stem, ext = (base[:dot], base[dot:]) if dot >= 0 else (base, "")
    return "%s_%06d%s" % (stem, step, ext)

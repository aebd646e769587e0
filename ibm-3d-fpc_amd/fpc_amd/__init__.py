"""fpc_amd -- Python binding of libfpc.so, the MI355X Struct-stencil stack and
time-step driver of the 3-D channel-with-cylinder solver.

The product is the C ABI in include/fpc_struct.h, include/fpc_dense.h and
include/fpc_sim.h; this module loads libfpc.so with ctypes and mirrors the
reference's host interface (JD63021/ibm-3d-fpc: Params and ParseArgs of
common.hpp:45-127, the Struct helpers of hypre_struct_utils.hpp:1-43, the HYPRE
solver calls of main.cu:173-221) for Python callers.  Device memory is passed
as raw pointers, e.g. ``tensor.data_ptr()`` of a torch CUDA tensor.

Nothing here falls back to a CPU implementation: if libfpc.so is missing or a
call fails, an exception is raised.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
REPO_ROOT = os.path.dirname(PKG_ROOT)
LIB_PATH = os.path.join(HERE, "libfpc.so")
CSRC = os.path.join(PKG_ROOT, "csrc")
INCLUDE = os.path.join(REPO_ROOT, "include")

ERROR_GENERIC, ERROR_MEMORY, ERROR_ARG, ERROR_CONV = 1, 2, 4, 256


class FpcError(RuntimeError):
    pass


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
    ("mg_precision", ctypes.c_int), ("vel_print", ctypes.c_int), ("p_print", ctypes.c_int),
    ("write_vtk", ctypes.c_int), ("vtk_every", ctypes.c_int), ("device", ctypes.c_int),
    ("vtk_filename", ctypes.c_char * 256),
]


class Params(ctypes.Structure):
    _fields_ = PARAM_FIELDS

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in PARAM_FIELDS}


class FlowOps(ctypes.Structure):
    """fpc_flow_ops (include/fpc_dense.h): a channel's operator data for the fused stages,
    device pointers."""
    _fields_ = ([(n, ctypes.c_int) for n in ("nx", "ny", "nz")]
                + [(n, ctypes.c_double) for n in ("dx", "dy", "dz", "rho", "dt")]
                + [(n, ctypes.c_void_p) for n in ("active_u", "active_v", "active_w", "fluid_p", "active_qu",
                                                   "active_qv", "active_qw", "scale_u", "scale_v", "scale_w",
                                                   "rhs_u_bc", "dcoef", "div_bc")]
                + [(n, ctypes.c_int) for n in ("n_fu", "n_fv", "n_px", "n_py")]
                + [(n, ctypes.c_void_p) for n in ("fu_idx", "fv_idx", "px_idx", "py_idx", "fu_w", "fv_w", "px_w",
                                                   "py_w")])


def build(jobs=8):
    """Compile libfpc.so for gfx950 with the package Makefile (hipcc + g++)."""
    subprocess.run(["make", "-s", "-C", CSRC, f"-j{jobs}"], check=True)


_lib = None

# (name, restype, argtypes) of every exported symbol declared in include/*.h
_vp, _cp, _ll, _d, _i = ctypes.c_void_p, ctypes.c_char_p, ctypes.c_longlong, ctypes.c_double, ctypes.c_int
_pi, _pd, _pvp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_void_p)
_sz = ctypes.c_size_t
_PP = ctypes.POINTER(Params)
SOLVER_FN = ctypes.CFUNCTYPE(ctypes.c_int, _vp, _vp, _vp, _vp)

SIGNATURES = {
    # fpc_struct.h
    "fpc_Initialize": (_i, []), "fpc_Finalize": (_i, []),
    "fpc_SetMemoryLocationDevice": (_i, []), "fpc_SetExecutionPolicyDevice": (_i, []),
    "fpc_GetError": (_i, []), "fpc_ClearError": (_i, [_i]), "fpc_GetErrorMessage": (_cp, []),
    "fpc_SetStream": (_i, [_vp]), "fpc_GetStream": (_vp, []),
    "fpc_StructGridCreate": (_i, [_i, _pvp]), "fpc_StructGridSetExtents": (_i, [_vp, _pi, _pi]),
    "fpc_StructGridAssemble": (_i, [_vp]), "fpc_StructGridDestroy": (_i, [_vp]),
    "fpc_StructStencilCreate": (_i, [_i, _i, _pvp]), "fpc_StructStencilSetElement": (_i, [_vp, _i, _pi]),
    "fpc_StructStencilDestroy": (_i, [_vp]),
    "fpc_StructMatrixCreate": (_i, [_vp, _vp, _pvp]), "fpc_StructMatrixInitialize": (_i, [_vp]),
    "fpc_StructMatrixSetBoxValues": (_i, [_vp, _pi, _pi, _i, _pi, _vp]),
    "fpc_StructMatrixGetBoxValues": (_i, [_vp, _pi, _pi, _i, _pi, _vp]),
    "fpc_StructMatrixAssemble": (_i, [_vp]), "fpc_StructMatrixMatvec": (_i, [_d, _vp, _vp, _d, _vp]),
    "fpc_StructMatrixDestroy": (_i, [_vp]),
    "fpc_StructVectorCreate": (_i, [_vp, _pvp]), "fpc_StructVectorInitialize": (_i, [_vp]),
    "fpc_StructVectorSetBoxValues": (_i, [_vp, _pi, _pi, _vp]),
    "fpc_StructVectorGetBoxValues": (_i, [_vp, _pi, _pi, _vp]),
    "fpc_StructVectorSetConstantValues": (_i, [_vp, _d]), "fpc_StructVectorAssemble": (_i, [_vp]),
    "fpc_StructVectorGetDataPointer": (_i, [_vp, _pvp]), "fpc_StructVectorDestroy": (_i, [_vp]),
    "fpc_StructGMRESCreate": (_i, [_pvp]), "fpc_StructFlexGMRESCreate": (_i, [_pvp]),
    "fpc_StructGMRESSetTol": (_i, [_vp, _d]), "fpc_StructGMRESSetAbsoluteTol": (_i, [_vp, _d]),
    "fpc_StructGMRESSetMaxIter": (_i, [_vp, _i]), "fpc_StructGMRESSetKDim": (_i, [_vp, _i]),
    "fpc_StructGMRESSetPrintLevel": (_i, [_vp, _i]), "fpc_StructGMRESSetLogging": (_i, [_vp, _i]),
    "fpc_StructGMRESSetPrecond": (_i, [_vp, _vp, _vp, _vp]),
    "fpc_StructGMRESSetup": (_i, [_vp, _vp, _vp, _vp]), "fpc_StructGMRESSolve": (_i, [_vp, _vp, _vp, _vp]),
    "fpc_StructGMRESGetNumIterations": (_i, [_vp, _pi]),
    "fpc_StructGMRESGetFinalRelativeResidualNorm": (_i, [_vp, _pd]), "fpc_StructGMRESDestroy": (_i, [_vp]),
    "fpc_StructJacobiCreate": (_i, [_pvp]), "fpc_StructJacobiSetMaxIter": (_i, [_vp, _i]),
    "fpc_StructJacobiSetTol": (_i, [_vp, _d]), "fpc_StructJacobiSetZeroGuess": (_i, [_vp]),
    "fpc_StructJacobiSetPrecision": (_i, [_vp, _i]),
    "fpc_StructJacobiSetNonZeroGuess": (_i, [_vp]), "fpc_StructJacobiSetup": (_i, [_vp, _vp, _vp, _vp]),
    "fpc_StructJacobiSolve": (_i, [_vp, _vp, _vp, _vp]), "fpc_StructJacobiDestroy": (_i, [_vp]),
    "fpc_StructPFMGCreate": (_i, [_pvp]), "fpc_StructPFMGSetMaxIter": (_i, [_vp, _i]),
    "fpc_StructPFMGSetTol": (_i, [_vp, _d]), "fpc_StructPFMGSetRelaxType": (_i, [_vp, _i]),
    "fpc_StructPFMGSetRAPType": (_i, [_vp, _i]), "fpc_StructPFMGSetNumPreRelax": (_i, [_vp, _i]),
    "fpc_StructPFMGSetNumPostRelax": (_i, [_vp, _i]), "fpc_StructPFMGSetMaxLevels": (_i, [_vp, _i]),
    "fpc_StructPFMGSetPrintLevel": (_i, [_vp, _i]), "fpc_StructPFMGSetZeroGuess": (_i, [_vp]),
    "fpc_StructPFMGSetNonZeroGuess": (_i, [_vp]), "fpc_StructPFMGSetPrecision": (_i, [_vp, _i]),
    "fpc_StructPFMGSetCoarseSize": (_i, [_vp, _i]),
    "fpc_StructMatrixGetExtruded": (_i, [_vp, _pi, ctypes.POINTER(ctypes.c_longlong)]),
    "fpc_StructPFMGGetLevelExtruded": (_i, [_vp, _i, _pi, ctypes.POINTER(ctypes.c_longlong)]),
    "fpc_StructPFMGSetup": (_i, [_vp, _vp, _vp, _vp]), "fpc_StructPFMGSolve": (_i, [_vp, _vp, _vp, _vp]),
    "fpc_StructPFMGGetNumIterations": (_i, [_vp, _pi]), "fpc_StructPFMGGetNumLevels": (_i, [_vp, _pi]),
    "fpc_StructPFMGGetLevelInfo": (_i, [_vp, _i, _pi, _pi, _pd]), "fpc_StructPFMGDestroy": (_i, [_vp]),
    "fpc_StructPFMGGetLevelMatrix": (_i, [_vp, _i, _vp]),
    "fpc_CreateGrid3D": (_vp, [_i, _i, _i]), "fpc_CreateStencil3D": (_vp, []),
    "fpc_SetMatrix3D": (_i, [_vp, _i, _i, _i, _vp]), "fpc_CreateMatrix3D": (_vp, [_vp, _vp, _i, _i, _i, _vp]),
    "fpc_SetVector3D": (_i, [_vp, _i, _i, _i, _vp]), "fpc_GetVector3D": (_i, [_vp, _i, _i, _i, _vp]),
    "fpc_CreateVector3D": (_vp, [_vp]),
    "fpc_MallocManagedC": (_i, [_pvp, _sz]), "fpc_MallocManagedD": (_i, [_pvp, _sz]),
    "fpc_MallocManagedI": (_i, [_pvp, _sz]), "fpc_MallocManagedB": (_i, [_pvp, _sz]),
    "fpc_CopyVecToManagedC": (_i, [_pvp, _vp, _sz]), "fpc_CopyVecToManagedD": (_i, [_pvp, _vp, _sz]),
    "fpc_CopyVecToManagedI": (_i, [_pvp, _vp, _sz]), "fpc_CopyVecToManagedB": (_i, [_pvp, _vp, _sz]),
    "fpc_Free": (_i, [_vp]), "fpc_MemcpyToHost": (_i, [_vp, _vp, _sz]), "fpc_MemcpyToDevice": (_i, [_vp, _vp, _sz]),
    "fpc_DeviceSynchronize": (_i, []),
    "fpc_SetTuning": (_i, [_cp, ctypes.c_longlong]),
    "fpc_ProfileEnable": (_i, [_cp]), "fpc_ProfileSample": (_i, [_i]), "fpc_ProfileReset": (_i, []),
    "fpc_ProfileReport": (_i, [_cp, _i]),
    # fpc_dense.h
    "fpc_apply_bc_cleanup": (_i, [_vp] * 7 + [_i] * 3),
    "fpc_conv_u": (_i, [_vp] * 5 + [_i] * 3 + [_d] * 3), "fpc_conv_v": (_i, [_vp] * 5 + [_i] * 3 + [_d] * 3),
    "fpc_conv_w": (_i, [_vp] * 5 + [_i] * 3 + [_d] * 3),
    "fpc_build_u_full": (_i, [_vp] * 9 + [_i] * 3 + [_d, _d, _i]),
    "fpc_build_v_full": (_i, [_vp] * 9 + [_i] * 3 + [_d, _d, _i]),
    "fpc_build_w_full": (_i, [_vp] * 9 + [_i] * 3 + [_d, _d, _i]),
    "fpc_scatter_u": (_i, [_vp] * 3 + [_i] * 3), "fpc_scatter_v": (_i, [_vp] * 3 + [_i] * 3),
    "fpc_scatter_w": (_i, [_vp] * 3 + [_i] * 3),
    "fpc_spmv_csr_1based": (_i, [_i] + [_vp] * 5),
    "fpc_pack_pressure_active": (_i, [_vp] * 4 + [_i]),
    "fpc_unpack_gp_u": (_i, [_vp] * 3 + [_i] * 3), "fpc_unpack_gp_v": (_i, [_vp] * 3 + [_i] * 4),
    "fpc_unpack_gp_w": (_i, [_vp] * 3 + [_i] * 5),
    "fpc_pack_projection_q_u": (_i, [_vp] * 3 + [_i] * 3), "fpc_pack_projection_q_v": (_i, [_vp] * 3 + [_i] * 4),
    "fpc_pack_projection_q_w": (_i, [_vp] * 3 + [_i] * 5),
    "fpc_build_pressure_rhs_full": (_i, [_vp] * 6 + [_i, _d, _d, _d]),
    "fpc_pack_phi_active": (_i, [_vp] * 4 + [_i]), "fpc_apply_q_correction": (_i, [_vp, _vp, _d, _i]),
    "fpc_unpack_projection_q_u": (_i, [_vp] * 3 + [_i] * 3),
    "fpc_unpack_projection_q_v": (_i, [_vp] * 3 + [_i] * 4),
    "fpc_unpack_projection_q_w": (_i, [_vp] * 3 + [_i] * 5),
    "fpc_add_pressure_phi": (_i, [_vp] * 4 + [_i]), "fpc_zero_doubles": (_i, [_vp, _ll]),
    "fpc_stage_explicit": (_i, [_vp] * 12 + [_i]), "fpc_stage_scatter": (_i, [_vp] * 11),
    "fpc_stage_divergence": (_i, [_vp] * 4 + [_d, _vp, _vp]), "fpc_stage_projection": (_i, [_vp] * 11),
    "fpc_stage_diagnostics": (_i, [_vp] * 6 + [_d, _i, _vp]),
    "fpc_stage_diagnostics_noalloc": (_i, [_vp] * 6 + [_d, _i, _vp]), "fpc_any_nonfinite": (_i, [_vp, _ll, _pi]),
    # fpc_sim.h
    "fpc_default_params": (None, [_PP]), "fpc_parse_args": (_i, [_i, ctypes.POINTER(_cp), _PP]),
    "fpc_setup_create": (_i, [_PP, _pvp]), "fpc_setup_len": (_ll, [_vp, _cp]), "fpc_setup_elsize": (_i, [_vp, _cp]),
    "fpc_setup_copy": (_i, [_vp, _cp, _vp, _ll]), "fpc_setup_int": (_ll, [_vp, _cp]), "fpc_setup_destroy": (_i, [_vp]),
    "fpc_sim_create": (_i, [_PP, _pvp]), "fpc_sim_step": (_i, [_vp, _i]), "fpc_sim_phase": (_i, [_vp, _i]),
    "fpc_sim_len": (_ll, [_vp, _cp]), "fpc_sim_copy": (_i, [_vp, _cp, _vp, _ll]),
    "fpc_sim_set": (_i, [_vp, _cp, _vp, _ll]), "fpc_sim_device_ptr": (_i, [_vp, _cp, _pvp]),
    "fpc_sim_scalar": (_d, [_vp, _cp]), "fpc_sim_timer": (_d, [_vp, _cp]), "fpc_sim_destroy": (_i, [_vp]),
    "fpc_write_vtk_cellcentered": (_i, [_cp] + [_vp] * 5 + [_i] * 3 + [_d] * 3),
    "fpc_vtk_filename_with_step": (_i, [_cp, _i, _cp, _i]),
    "fpc_write_fields": (_i, [_cp] + [_i] * 4 + [_vp] * 4),
    "fpc_sim_write_vtk": (_i, [_vp, _cp]), "fpc_sim_write_fields": (_i, [_vp, _cp]),
    "fpc_sim_stream": (_vp, [_vp]), "fpc_sim_flow_ops": (_i, [_vp, ctypes.POINTER(FlowOps)]),
    "fpc_ensemble_create": (_i, [_PP, _i, _i, _pvp]), "fpc_ensemble_step": (_i, [_vp, _i]),
    "fpc_ensemble_size": (_i, [_vp]), "fpc_ensemble_member": (_vp, [_vp, _i]),
    "fpc_ensemble_tiles": (_i, [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_longlong)]),
    "fpc_ensemble_offsets": (_i, [_vp, _cp, ctypes.POINTER(ctypes.c_longlong)]),
    "fpc_ensemble_pack": (_i, [_vp, _cp, _vp, _ll]), "fpc_ensemble_destroy": (_i, [_vp]),
    "fpc_adjoint_create": (_i, [_vp, _pvp]),
    "fpc_adjoint_step_vjp": (_i, [_vp, _i] + [_vp] * 9 + [ctypes.POINTER(ctypes.c_double)]),
    "fpc_adjoint_drag_gradient": (_i, [_vp, _i, _i, _i, ctypes.POINTER(ctypes.c_double), _vp,
                                      ctypes.POINTER(ctypes.c_double)]),
    "fpc_adjoint_destroy": (_i, [_vp]),
}


def lib():
    """Load libfpc.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FpcError(f"libfpc.so not built at {LIB_PATH}; run fpc_amd.build() or __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc:
        msg = lib().fpc_GetErrorMessage()
        raise FpcError(f"{what} failed with code {rc}: {msg.decode() if msg else ''}")
    return rc


def default_params(**kw):
    """Params with the reference defaults (common.hpp:45-85), then overrides."""
    p = Params()
    lib().fpc_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def parse_args(argv, base=None):
    """ParseArgs with the reference flag names (common.hpp:87-127)."""
    p = base if base is not None else default_params()
    args = ["prog"] + list(argv)
    arr = (ctypes.c_char_p * len(args))(*[a.encode() for a in args])
    check(lib().fpc_parse_args(len(args), arr, ctypes.byref(p)), "fpc_parse_args")
    return p


_NP = {1: np.uint8, 4: np.int32, 8: np.float64}
_INT_ARRAYS = ("id_", "D_rowptr", "D_colind", "G_rowptr", "G_colind")


class Setup:
    """Host setup arrays (masks, ids, embedded boxes, CSR) built by libfpc."""

    INTS = ("Nu", "Nv", "Nw", "Np", "Nqu", "Nqv", "Nqw", "n_pruned")

    def __init__(self, params):
        self.params = params
        h = ctypes.c_void_p()
        check(lib().fpc_setup_create(ctypes.byref(params), ctypes.byref(h)), "fpc_setup_create")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.fpc_setup_destroy(self.h)
            self.h = None

    def get(self, name):
        n = lib().fpc_setup_len(self.h, name.encode())
        if n < 0:
            raise KeyError(name)
        es = lib().fpc_setup_elsize(self.h, name.encode())
        dt = np.int32 if name.startswith(_INT_ARRAYS) else _NP[es]
        out = np.empty(n, dtype=dt)
        check(lib().fpc_setup_copy(self.h, name.encode(), out.ctypes.data, out.nbytes), f"setup copy {name}")
        return out

    def count(self, name):
        return int(lib().fpc_setup_int(self.h, name.encode()))


class Sim:
    """The time-step driver (main.cu:249-376) on the GPU."""

    SCALARS = ("step", "ramp", "div_l2", "umax", "vmax", "wmax", "its_u", "its_v", "its_w", "its_p",
               "relres_u", "relres_v", "relres_w", "relres_p", "fx", "fy", "cd", "cl",
               "err_u", "err_v", "err_w", "err_p")
    FIELDS = ("u", "v", "w", "p", "u_tilde", "v_tilde", "w_tilde", "rhs_u", "x_u", "rhs_v", "x_v", "rhs_w",
              "x_w", "rhs_p", "x_p", "uin_step", "qout")

    def __init__(self, params):
        self.params = params
        h = ctypes.c_void_p()
        check(lib().fpc_sim_create(ctypes.byref(params), ctypes.byref(h)), "fpc_sim_create")
        self.h = h
        self._owned = True

    def close(self):
        # ensemble members belong to their ensemble
        if getattr(self, "h", None) and getattr(self, "_owned", True) and _lib is not None:
            _lib.fpc_sim_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()

    def step(self, n=1):
        check(lib().fpc_sim_step(self.h, n), "fpc_sim_step")

    def phase(self, ph):
        check(lib().fpc_sim_phase(self.h, ph), f"fpc_sim_phase({ph})")

    def get(self, name):
        n = lib().fpc_sim_len(self.h, name.encode())
        if n < 0:
            raise KeyError(name)
        if name in self.FIELDS:
            dt = np.float64
        else:
            es = lib().fpc_setup_elsize  # arrays of the setup: element size from the name
            dt = np.int32 if name.startswith(_INT_ARRAYS) else (np.uint8 if name.startswith(("active", "fluid")) else np.float64)
        out = np.empty(n, dtype=dt)
        check(lib().fpc_sim_copy(self.h, name.encode(), out.ctypes.data, out.nbytes), f"sim copy {name}")
        return out

    def set(self, name, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        check(lib().fpc_sim_set(self.h, name.encode(), arr.ctypes.data, arr.nbytes), f"sim set {name}")

    def set_step(self, n):
        v = ctypes.c_int(int(n))
        check(lib().fpc_sim_set(self.h, b"step", ctypes.addressof(v), ctypes.sizeof(v)), "sim set step")

    def device_ptr(self, name):
        p = ctypes.c_void_p()
        check(lib().fpc_sim_device_ptr(self.h, name.encode(), ctypes.byref(p)), f"device_ptr {name}")
        return p.value

    def scalar(self, name):
        return float(lib().fpc_sim_scalar(self.h, name.encode()))

    def scalars(self):
        return {k: self.scalar(k) for k in self.SCALARS}

    def timer(self, name):
        return float(lib().fpc_sim_timer(self.h, name.encode()))

    def write_vtk(self, filename):
        """The reference driver's cell-centred VTK file of the current fields."""
        check(lib().fpc_sim_write_vtk(self.h, os.fsencode(filename)), "fpc_sim_write_vtk")

    def write_fields(self, path):
        """Raw dump of u, v, w, p (read_fields reads it back)."""
        check(lib().fpc_sim_write_fields(self.h, os.fsencode(path)), "fpc_sim_write_fields")

    def flow_ops(self):
        """The operator data of the fused stages (FlowOps, device pointers owned by the sim)."""
        o = FlowOps()
        check(lib().fpc_sim_flow_ops(self.h, ctypes.byref(o)), "fpc_sim_flow_ops")
        return o

    def stream(self):
        """The simulation's HIP stream (its calls run there)."""
        return lib().fpc_sim_stream(self.h)


def vtk_filename_with_step(base, step):
    """base with "_%06d" (step) before its extension (solver_helpers.hpp:470-482)."""
    out = ctypes.create_string_buffer(len(os.fsencode(base)) + 64)
    check(lib().fpc_vtk_filename_with_step(os.fsencode(base), int(step), out, len(out)), "vtk_filename_with_step")
    return os.fsdecode(out.value)


def write_vtk_cellcentered(filename, u, v, w, p, fluid_p, nx, ny, nz, Lx, Ly, Lz):
    """The reference's cell-centred VTK file from host arrays (byte-identical format)."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, w, p)]
    fl = np.ascontiguousarray(fluid_p, dtype=np.uint8)
    check(lib().fpc_write_vtk_cellcentered(os.fsencode(filename), *[a.ctypes.data for a in arrs], fl.ctypes.data,
                                           nx, ny, nz, Lx, Ly, Lz), "write_vtk_cellcentered")


def write_fields(path, step, nx, ny, nz, u, v, w, p):
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, w, p)]
    check(lib().fpc_write_fields(os.fsencode(path), step, nx, ny, nz, *[a.ctypes.data for a in arrs]), "write_fields")


def read_fields(path):
    """A raw field dump (fpc_write_fields): dict with nx, ny, nz, step, u, v, w, p."""
    with open(path, "rb") as f:
        head = f.read(40)
        if head[:8] != b"FPCFLD01":
            raise FpcError(f"{path}: not a field dump")
        nx, ny, nz, step = np.frombuffer(head[8:24], dtype="<i4")
        total = int(np.frombuffer(head[24:32], dtype="<i8")[0])
        data = np.frombuffer(f.read(8 * total), dtype="<f8")
    nx, ny, nz = int(nx), int(ny), int(nz)
    sizes = [(nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1), nx * ny * nz]
    if sum(sizes) != total or data.size != total:
        raise FpcError(f"{path}: truncated or inconsistent dump")
    out = {"nx": nx, "ny": ny, "nz": nz, "step": int(step)}
    o = 0
    for name, n in zip("uvwp", sizes):
        out[name] = data[o:o + n].copy()
        o += n
    return out


def set_tuning(key, value):
    """Launch tuning of the z-marching stencil kernels ("tile_min_points", "stencil_kz")."""
    check(lib().fpc_SetTuning(key.encode(), int(value)), f"SetTuning({key})")


class Adjoint:
    """Discrete adjoint of a simulation's time step (fpc_adjoint_*): gradients of the
    time-averaged drag coefficient with respect to the inlet amplitude field theta
    (ny*nz) and nu."""

    def __init__(self, sim):
        self.sim = sim
        self.h = ctypes.c_void_p()
        check(lib().fpc_adjoint_create(sim.h, ctypes.byref(self.h)), "adjoint create")

    def drag_gradient(self, nsteps, avg_from, every):
        """(J, dJ/dtheta [ny*nz], dJ/dnu) for J = mean c_D over steps avg_from..nsteps."""
        p = self.sim.params
        g = np.zeros(p.ny * p.nz)
        J, gnu = ctypes.c_double(), ctypes.c_double()
        check(lib().fpc_adjoint_drag_gradient(self.h, nsteps, avg_from, every, ctypes.byref(J), g.ctypes.data,
                                              ctypes.byref(gnu)), "drag gradient")
        return J.value, g, gnu.value

    def step_vjp(self, step, state, adj, theta_bar):
        """One step's VJP on torch CUDA tensors: state = (u, v, w, p) before step `step`;
        adj = (au, av, aw, ap) adjoint after the step, overwritten with the adjoint
        before it; theta_bar (ny*nz) accumulated.  Returns the nu adjoint."""
        nu = ctypes.c_double(0.0)
        ptr = lambda t: ctypes.c_void_p(t.data_ptr())
        check(lib().fpc_adjoint_step_vjp(self.h, step, *[ptr(t) for t in state], *[ptr(t) for t in adj],
                                         ptr(theta_bar), ctypes.byref(nu)), "step vjp")
        return nu.value

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.fpc_adjoint_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()


def ensemble_workload(n=64, seed=0, nx=(64, 256), nyz=(16, 48), xc=(0.4, 0.7), yc=(0.15, 0.26), re=(10.0, 100.0),
                      **overrides):
    """Seeded uneven channel instances: nx in [nx0, nx1], ny = nz in [nyz0, nyz1],
    cylinder (r = 0.05) centre x in xc, y in yc, Re = (4 Um / 9) D / nu in re with
    nu = 1e-3; dt = h_min / (8 Um); inlet profile + 1e-3 Um noise; solvers as the
    production configuration (GMRES(30), rtol 1e-8)."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(n):
        gx = int(rng.integers(nx[0], nx[1] + 1))
        g = int(rng.integers(nyz[0], nyz[1] + 1))
        Re = float(rng.uniform(*re))
        um = Re * 1e-3 / (4.0 / 9.0 * 0.1)
        h = min(2.5 / gx, 0.41 / g)
        kw = dict(nx=gx, ny=g, nz=g, Lx=2.5, Ly=0.41, Lz=0.41, xcirc=float(rng.uniform(*xc)),
                  ycirc=float(rng.uniform(*yc)), Rcirc=0.05, Um=um, nu=1e-3, rho=1.0, dt=h / (8.0 * um),
                  vel_kdim=30, p_kdim=30, vel_tol=1e-8, p_tol=1e-8, vel_maxit=250, p_maxit=250, init_profile=1,
                  noise_amp=1e-3 * um, seed=seed * 1000 + q)
        kw.update(overrides)
        out.append(default_params(**kw))
    return out


class Ensemble:
    """Independent channels (fpc_ensemble_*): stepped concurrently, one HIP stream per
    instance and a pool of `nthreads` host worker threads, or, with nthreads=0,
    batched on one stream with one launch per stage over every instance."""

    def __init__(self, params_list, nthreads=8):
        self.params = list(params_list)
        arr = (Params * len(self.params))(*self.params)
        self.h = ctypes.c_void_p()
        check(lib().fpc_ensemble_create(arr, len(self.params), int(nthreads), ctypes.byref(self.h)), "ensemble create")
        self.members = []
        for i in range(len(self.params)):
            m = Sim.__new__(Sim)
            m.h = ctypes.c_void_p(lib().fpc_ensemble_member(self.h, i))
            m.params = self.params[i]
            m._owned = False
            self.members.append(m)

    def step(self, n=1):
        check(lib().fpc_ensemble_step(self.h, n), "ensemble step")

    def tiles(self):
        """(z-chunk planes, smallest tiled box) picked by the autotuner of a batched ensemble."""
        kz, tmin = ctypes.c_int(), ctypes.c_longlong()
        check(lib().fpc_ensemble_tiles(self.h, ctypes.byref(kz), ctypes.byref(tmin)), "ensemble tiles")
        return kz.value, tmin.value

    def offsets(self, field):
        off = (ctypes.c_longlong * (len(self.params) + 1))()
        check(lib().fpc_ensemble_offsets(self.h, field.encode(), off), "ensemble offsets")
        return np.array(off[:], dtype=np.int64)

    def pack(self, field):
        off = self.offsets(field)
        out = np.empty(int(off[-1]), dtype=np.float64)
        check(lib().fpc_ensemble_pack(self.h, field.encode(), out.ctypes.data, out.nbytes), "ensemble pack")
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().fpc_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def profile_sample(every=1):
    """Time only every `every`-th launch of the enabled classes."""
    check(lib().fpc_ProfileSample(int(every)), "ProfileSample")


def profile_enable(names="*"):
    check(lib().fpc_ProfileEnable(names.encode() if names else None), "ProfileEnable")
    lib().fpc_ProfileReset()


def profile_report():
    """{kernel class: (launches, total_ms, algorithmic_bytes)} of the timed launches."""
    n = lib().fpc_ProfileReport(None, 0)
    buf = ctypes.create_string_buffer(max(n, 1))
    lib().fpc_ProfileReport(buf, n)
    out = {}
    for line in buf.value.decode().splitlines():
        name, cnt, ms, by = line.split()
        out[name] = (int(cnt), float(ms), float(by))
    return out


def exported_symbols():
    """Names declared in include/*.h (parsed from the headers)."""
    import re
    names = set()
    for fn in sorted(os.listdir(INCLUDE)):
        if not fn.endswith(".h"):
            continue
        txt = open(os.path.join(INCLUDE, fn)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(fpc_[A-Za-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    names.discard("fpc_PtrToStructSolverFcn")
    return sorted(names)


# ---------------------------------------------------------------- Struct helpers
def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if isinstance(a, np.ndarray) else ctypes.c_void_p(int(a))


class StructBox:
    """A box grid, the 7-point stencil and one matrix (CreateGrid3D / CreateStencil3D /
    CreateMatrix3D of hypre_struct_utils.hpp:1-20).  vals: AoS, 7 per point (host or device)."""

    def __init__(self, nx, ny, nz, vals):
        L = lib()
        self.n = (nx, ny, nz)
        self.N = nx * ny * nz
        self.grid = L.fpc_CreateGrid3D(nx, ny, nz)
        self.stencil = L.fpc_CreateStencil3D()
        if not self.grid or not self.stencil:
            check(FpcError and 1, "CreateGrid3D/CreateStencil3D")
        self._vals = np.ascontiguousarray(vals, dtype=np.float64) if isinstance(vals, np.ndarray) else vals
        self.A = L.fpc_CreateMatrix3D(self.grid, self.stencil, nx, ny, nz, _ptr(self._vals))
        if not self.A:
            check(1, "CreateMatrix3D")

    def vector(self, values=None):
        v = StructVec(self)
        if values is not None:
            v.set(values)
        return v

    def extruded(self):
        """(wall planes m, flagged points) of the copy the solve loops read; m = 0: full planes."""
        m, f = ctypes.c_int(), ctypes.c_longlong()
        check(lib().fpc_StructMatrixGetExtruded(self.A, ctypes.byref(m), ctypes.byref(f)), "GetExtruded")
        return m.value, f.value

    def matvec(self, x, alpha=1.0, beta=0.0, y=None):
        xv = self.vector(x)
        yv = self.vector(y if y is not None else np.zeros(self.N))
        check(lib().fpc_StructMatrixMatvec(alpha, self.A, xv.h, beta, yv.h), "Matvec")
        return yv.get()

    def close(self):
        L = _lib
        if L is None:
            return
        if getattr(self, "A", None):
            L.fpc_StructMatrixDestroy(self.A)
            self.A = None
        if getattr(self, "stencil", None):
            L.fpc_StructStencilDestroy(self.stencil)
            self.stencil = None
        if getattr(self, "grid", None):
            L.fpc_StructGridDestroy(self.grid)
            self.grid = None

    def __del__(self):
        self.close()


class StructVec:
    def __init__(self, box):
        self.box = box
        self.h = lib().fpc_CreateVector3D(box.grid)
        if not self.h:
            check(1, "CreateVector3D")

    def set(self, values):
        a = np.ascontiguousarray(values, dtype=np.float64)
        assert a.size == self.box.N
        nx, ny, nz = self.box.n
        check(lib().fpc_SetVector3D(self.h, nx, ny, nz, _ptr(a)), "SetVector3D")

    def get(self):
        out = np.empty(self.box.N)
        nx, ny, nz = self.box.n
        check(lib().fpc_GetVector3D(self.h, nx, ny, nz, _ptr(out)), "GetVector3D")
        return out

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.fpc_StructVectorDestroy(self.h)
            self.h = None


class Solver:
    """GMRES (or flexible GMRES) with a Jacobi / PFMG / no preconditioner, set up as main.cu:173-221."""

    def __init__(self, box, tol=1e-8, maxit=250, kdim=30, precond="jacobi", pc_iters=1, relax_type=1,
                 rap_type=1, flexible=False, mg_bits=64, jacobi_bits=64, coarse_size=0):
        L = lib()
        self.box = box
        self.s = ctypes.c_void_p()
        check((L.fpc_StructFlexGMRESCreate if flexible else L.fpc_StructGMRESCreate)(ctypes.byref(self.s)), "create")
        check(L.fpc_StructGMRESSetTol(self.s, tol), "tol")
        check(L.fpc_StructGMRESSetMaxIter(self.s, maxit), "maxit")
        check(L.fpc_StructGMRESSetKDim(self.s, kdim), "kdim")
        self.pc = ctypes.c_void_p()
        self.kind = precond
        if precond == "jacobi":
            check(L.fpc_StructJacobiCreate(ctypes.byref(self.pc)), "jacobi")
            check(L.fpc_StructJacobiSetMaxIter(self.pc, pc_iters), "jacobi maxit")
            check(L.fpc_StructJacobiSetZeroGuess(self.pc), "zero guess")
            check(L.fpc_StructJacobiSetPrecision(self.pc, jacobi_bits), "jacobi precision")
            check(L.fpc_StructGMRESSetPrecond(self.s, ctypes.cast(L.fpc_StructJacobiSolve, ctypes.c_void_p),
                                              ctypes.cast(L.fpc_StructJacobiSetup, ctypes.c_void_p), self.pc), "pc")
        elif precond == "pfmg":
            check(L.fpc_StructPFMGCreate(ctypes.byref(self.pc)), "pfmg")
            check(L.fpc_StructPFMGSetMaxIter(self.pc, pc_iters), "pfmg maxit")
            check(L.fpc_StructPFMGSetTol(self.pc, 0.0), "pfmg tol")
            check(L.fpc_StructPFMGSetRelaxType(self.pc, relax_type), "relax")
            check(L.fpc_StructPFMGSetRAPType(self.pc, rap_type), "rap")
            check(L.fpc_StructPFMGSetZeroGuess(self.pc), "zero guess")
            check(L.fpc_StructPFMGSetPrecision(self.pc, mg_bits), "precision")
            check(L.fpc_StructPFMGSetCoarseSize(self.pc, coarse_size), "coarse size")
            check(L.fpc_StructGMRESSetPrecond(self.s, ctypes.cast(L.fpc_StructPFMGSolve, ctypes.c_void_p),
                                              ctypes.cast(L.fpc_StructPFMGSetup, ctypes.c_void_p), self.pc), "pc")
        self.b = box.vector(np.zeros(box.N))
        self.x = box.vector(np.zeros(box.N))
        check(L.fpc_StructGMRESSetup(self.s, box.A, self.b.h, self.x.h), "GMRESSetup")

    def solve(self, b, x0=None):
        L = lib()
        self.b.set(b)
        self.x.set(np.zeros(self.box.N) if x0 is None else x0)
        rc = L.fpc_StructGMRESSolve(self.s, self.box.A, self.b.h, self.x.h)
        if rc & ~ERROR_CONV:
            check(rc, "GMRESSolve")
        if rc & ERROR_CONV:
            L.fpc_ClearError(ERROR_CONV)
        its = ctypes.c_int()
        rel = ctypes.c_double()
        L.fpc_StructGMRESGetNumIterations(self.s, ctypes.byref(its))
        L.fpc_StructGMRESGetFinalRelativeResidualNorm(self.s, ctypes.byref(rel))
        return self.x.get(), its.value, rel.value, rc

    def levels(self):
        L = lib()
        n = ctypes.c_int()
        check(L.fpc_StructPFMGGetNumLevels(self.pc, ctypes.byref(n)), "levels")
        out = []
        for l in range(n.value):
            dims = (ctypes.c_int * 3)()
            cd = ctypes.c_int()
            w = ctypes.c_double()
            check(L.fpc_StructPFMGGetLevelInfo(self.pc, l, dims, ctypes.byref(cd), ctypes.byref(w)), "level info")
            out.append((tuple(dims), cd.value, w.value))
        return out

    def levels_extruded(self):
        """Per MG level: (wall planes m, flagged points) of what its fused kernels read."""
        out = []
        for l in range(len(self.levels())):
            m, f = ctypes.c_int(), ctypes.c_longlong()
            check(lib().fpc_StructPFMGGetLevelExtruded(self.pc, l, ctypes.byref(m), ctypes.byref(f)), "level extruded")
            out.append((m.value, f.value))
        return out

    def level_matrix(self, level):
        dims = self.levels()[level][0]
        out = np.empty(7 * dims[0] * dims[1] * dims[2])
        check(lib().fpc_StructPFMGGetLevelMatrix(self.pc, level, _ptr(out)), "level matrix")
        return out

    def __del__(self):
        L = _lib
        if L is None:
            return
        if getattr(self, "s", None):
            L.fpc_StructGMRESDestroy(self.s)
            self.s = None
        if getattr(self, "pc", None):
            if self.kind == "jacobi":
                L.fpc_StructJacobiDestroy(self.pc)
            elif self.kind == "pfmg":
                L.fpc_StructPFMGDestroy(self.pc)
            self.pc = None


def pfmg_apply(box, b, relax_type=1, mg_bits=64, coarse_size=0):
    """One V-cycle from zero (HYPRE_StructPFMGSolve with MaxIter 1, zero guess)."""
    L = lib()
    pc = ctypes.c_void_p()
    check(L.fpc_StructPFMGCreate(ctypes.byref(pc)), "pfmg")
    try:
        check(L.fpc_StructPFMGSetMaxIter(pc, 1), "maxit")
        check(L.fpc_StructPFMGSetRelaxType(pc, relax_type), "relax")
        check(L.fpc_StructPFMGSetZeroGuess(pc), "zero")
        check(L.fpc_StructPFMGSetPrecision(pc, mg_bits), "bits")
        check(L.fpc_StructPFMGSetCoarseSize(pc, coarse_size), "coarse size")
        bv = box.vector(b)
        xv = box.vector(np.zeros(box.N))
        check(L.fpc_StructPFMGSetup(pc, box.A, bv.h, xv.h), "setup")
        check(L.fpc_StructPFMGSolve(pc, box.A, bv.h, xv.h), "solve")
        return xv.get()
    finally:
        L.fpc_StructPFMGDestroy(pc)

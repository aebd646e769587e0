"""The C ABI boundary without a GPU: libfpc.so loads, exports every function
declared in include/*.h, keeps the reference's Params defaults and flag names,
and rejects bad arguments with the HYPRE-style error codes."""
import ctypes
import os
import subprocess

import pytest

import fpc_amd as F
import oracle as O


def test_library_loads_and_exports_every_declared_symbol():
    F.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", F.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.split()[1:2] == ["T"]}
    declared = F.exported_symbols()
    assert len(declared) > 100
    missing = [n for n in declared if n not in exported]
    assert not missing, missing
    assert not [n for n in declared if n not in F.SIGNATURES]


def test_params_defaults_are_the_reference_defaults():
    p = F.default_params()
    o = O.default_params()
    # common.hpp:45-85
    want = dict(nx=32, ny=12, nz=12, Lx=2.2, Ly=0.41, Lz=0.41, rho=1.0, nu=1e-3, Um=0.45, dt=2e-3, nsteps=10,
                print_every=1, inlet_ramp_steps=200, use_zero_initial_guess=1, nsub_beta=10, beta_tol=1e-8,
                alpha_tol=1e-10, xcirc=0.2, ycirc=0.2, Rcirc=0.05, vel_maxit=250, vel_kdim=50, vel_tol=1e-6,
                vel_jacobi_iters=1, p_maxit=250, p_kdim=50, p_tol=1e-9, p_pc_maxit=1, p_pc_relax_type=1,
                p_pc_rap_type=1)
    for k, v in want.items():
        assert getattr(p, k) == v, k
        assert getattr(o, k) == v, k
    assert p.mg_precision == 64


def test_parse_args_uses_reference_flags():
    argv = ("-nx 128 -ny 32 -nz 32 -Lx 2.5 -xcirc 0.5 -ycirc 0.2 -Um 0.45 -nu 1e-3 -dt 1e-3 -nsteps 20 "
            "-vel-tol 1e-8 -p-tol 1e-8 -vel-kdim 30 -p-pc-relax-type 2 -write-vtk 0 -unknown 5 -mg-precision 32").split()
    p = F.parse_args(argv)
    assert (p.nx, p.ny, p.nz, p.nsteps, p.vel_kdim, p.p_pc_relax_type, p.mg_precision) == (128, 32, 32, 20, 30, 2, 32)
    assert (p.Lx, p.xcirc, p.dt, p.vel_tol, p.p_tol) == (2.5, 0.5, 1e-3, 1e-8, 1e-8)


def test_bad_arguments_return_arg_error():
    L = F.lib()
    g = ctypes.c_void_p()
    assert L.fpc_StructGridCreate(0, ctypes.byref(g)) == F.ERROR_ARG
    st = ctypes.c_void_p()
    assert L.fpc_StructStencilCreate(3, 9, ctypes.byref(st)) == F.ERROR_ARG
    h = ctypes.c_void_p()
    bad = F.default_params(nx=1)
    assert L.fpc_setup_create(ctypes.byref(bad), ctypes.byref(h)) == F.ERROR_ARG
    assert L.fpc_ClearError(F.ERROR_ARG) == 0


def test_tuning_keys_reject_bad_values():
    L = F.lib()
    for key, bad, good in (("mg_block", 100, 64), ("mg_block", 512, 256), ("p_flexible", 2, 1), ("mg_graph", -1, 0),
                           ("stencil_kz", -1, 0)):
        assert L.fpc_SetTuning(key.encode(), bad) == F.ERROR_ARG, (key, bad)
        assert L.fpc_ClearError(F.ERROR_ARG) == 0
        assert L.fpc_SetTuning(key.encode(), good) == 0, (key, good)
    assert L.fpc_SetTuning(b"no_such_key", 1) == F.ERROR_ARG
    assert L.fpc_ClearError(F.ERROR_ARG) == 0


def test_setup_runs_on_the_host_without_a_gpu():
    S = F.Setup(F.default_params())
    assert S.count("Np") > 0
    assert S.get("Ap_vals").size == 7 * 32 * 12 * 12


def test_product_module_has_no_oracle_dependency():
    src_dirs = [os.path.join(F.PKG_ROOT, "csrc"), F.HERE]
    for d in src_dirs:
        for fn in os.listdir(d):
            if fn.endswith((".py", ".hip", ".cpp", ".hpp", ".h")):
                txt = open(os.path.join(d, fn)).read()
                assert "oracle" not in txt.lower(), fn


REFERENCE_STYLE_CALLER = r"""
// The solver block of a HYPRE Struct caller (the shape of main.cu:151-221,
// 284-300 and 391-404), compiled against the compat header.
#include <cstdio>
#include <vector>
#include "fpc_hypre_compat.h"
#include "fpc_dense.h"
#include "fpc_sim.h"

int main() {
  const int nx = 8, ny = 4, nz = 4;
  HYPRE_Initialize();
  HYPRE_SetMemoryLocation(HYPRE_MEMORY_DEVICE);
  HYPRE_SetExecutionPolicy(HYPRE_EXEC_DEVICE);
  std::vector<double> vals(7 * nx * ny * nz, 0.0);
  for (size_t q = 0; q < vals.size(); q += 7) vals[q] = 1.0;
  HYPRE_Complex *vals_d = nullptr;
  CopyVecToManagedC(&vals_d, vals);
  HYPRE_StructGrid g = CreateGrid3D(nx, ny, nz);
  HYPRE_StructStencil st = CreateStencil3D();
  HYPRE_StructMatrix A = CreateMatrix3D(g, st, nx, ny, nz, vals_d);
  HYPRE_StructVector b = CreateVector3D(g), x = CreateVector3D(g);
  HYPRE_StructSolver s, pc;
  HYPRE_StructGMRESCreate(MPI_COMM_WORLD, &s);
  HYPRE_StructGMRESSetTol(s, 1e-8);
  HYPRE_StructGMRESSetKDim(s, 30);
  HYPRE_StructPFMGCreate(MPI_COMM_WORLD, &pc);
  HYPRE_StructPFMGSetMaxIter(pc, 1);
  HYPRE_StructPFMGSetZeroGuess(pc);
  HYPRE_StructGMRESSetPrecond(s, HYPRE_PtrToStructSolverFcn(HYPRE_StructPFMGSolve),
                              HYPRE_PtrToStructSolverFcn(HYPRE_StructPFMGSetup), pc);
  HYPRE_StructGMRESSetup(s, A, b, x);
  int rc = HYPRE_StructGMRESSolve(s, A, b, x);
  if (rc == HYPRE_ERROR_CONV) HYPRE_ClearError(HYPRE_ERROR_CONV);
  int its = 0;
  HYPRE_StructGMRESGetNumIterations(s, &its);
  fpc_params par;
  fpc_default_params(&par);
  std::printf("%d %d %d\n", rc, its, par.nx);
  HYPRE_StructGMRESDestroy(s);
  HYPRE_StructPFMGDestroy(pc);
  HYPRE_StructMatrixDestroy(A);
  HYPRE_Finalize();
  return 0;
}
"""


def test_reference_style_caller_compiles_and_links(tmp_path):
    # the drop-in boundary: a caller written against HYPRE Struct names builds
    # against include/ and links against libfpc.so (not run: no GPU here)
    F.lib()
    src = tmp_path / "caller.cpp"
    src.write_text(REFERENCE_STYLE_CALLER)
    exe = tmp_path / "caller"
    libdir = os.path.dirname(F.LIB_PATH)
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    cmd = ["g++", "-std=c++17", "-Wall", "-Werror", "-I", inc, str(src), "-o", str(exe), "-L", libdir, "-lfpc",
           "-Wl,-rpath-link,/opt/rocm/lib", "-Wl,-rpath," + libdir]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert exe.exists()

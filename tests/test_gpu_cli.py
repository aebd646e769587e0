"""The executables on the GPU, against the oracle.

* fpc_main, the command-line driver: the reference's flags, the device line, the
  banner (main.cu:223-245) with the oracle's setup counts, one line per printed
  step (main.cu:365-368), the timing summary (main.cu:380-389), and the VTK files
  at steps 1, nsteps and every vtk_every-th with the "_%06d" suffix
  (main.cu:370-378), byte for byte the oracle writer's file of the same fields.
* tests/caller/reference_caller, a program written like main.cu: HYPRE Struct
  names through fpc_hypre_compat.h and one fpc_dense.h call per reference kernel
  launch in the reference's order.  Its fields after the run equal the oracle's
  within 1e-9 (rtol 1e-12), with the same iteration counts per step.

Both executables are built by __graft_entry__.build() (package and
tests/caller Makefiles)."""
import os
import subprocess

import numpy as np
import pytest

import fpc_amd as F
import oracle as O
from helpers import rel_l2

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MAIN = os.path.join(REPO, "ibm-3d-fpc_amd", "fpc_amd", "fpc_main")
CALLER = os.path.join(REPO, "tests", "caller", "reference_caller")

FLAGS = dict(nx=32, ny=12, nz=12, Lx=2.2, xcirc=0.2, ycirc=0.2, Um=0.45, dt=2e-3, nsteps=4, inlet_ramp_steps=2,
             vel_tol=1e-12, p_tol=1e-12, vel_kdim=30, p_kdim=30)
FLAG_NAMES = {"vel_tol": "-vel-tol", "p_tol": "-p-tol", "vel_kdim": "-vel-kdim", "p_kdim": "-p-kdim",
              "inlet_ramp_steps": "-inlet-ramp-steps"}


def _argv(kw):
    out = []
    for k, v in kw.items():
        out += [FLAG_NAMES.get(k, "-" + k), repr(v) if isinstance(v, float) else str(v)]
    return out


def _run(cmd, timeout=120, cwd=None):
    assert os.path.exists(cmd[0]), f"{cmd[0]} not built: run __graft_entry__.build()"
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, cwd=cwd)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    return r.stdout


def _step_lines(out):
    rows = []
    for line in out.splitlines():
        parts = line.split()
        if len(parts) == 8 and parts[0].isdigit() and parts[6].startswith("("):
            its = [int(x) for x in parts[6].strip("()").split(",")]
            rel = [float(x) for x in parts[7].strip("()").split(",")]
            rows.append((int(parts[0]), float(parts[1]), [float(x) for x in parts[2:6]], its, rel))
    return rows


@pytest.fixture(scope="module")
def oracle_run():
    """The oracle's per-step scalars and fields for FLAGS (reference zero start)."""
    o = O.Sim(O.default_params(**{k: v for k, v in FLAGS.items() if k != "nsteps"}))
    rows = []
    for n in range(FLAGS["nsteps"]):
        o.step(1)
        s = o.scalars()
        rows.append((s, {f: o.get(f) for f in "uvwp"}))
    return rows


def test_driver_log_vtk_and_fields(torch_cuda, oracle_run, tmp_path):
    base = str(tmp_path / "chan.vtk")
    dump = str(tmp_path / "fields.bin")
    out = _run([MAIN] + _argv(FLAGS) + ["-write-vtk", "1", "-vtk-every", "3", "-vtk", base, "-dump-fields", dump])
    lines = out.splitlines()
    assert lines[0].startswith("GPU                   : ")
    S = O.Setup(O.default_params(**{k: v for k, v in FLAGS.items() if k != "nsteps"}))
    c = {k: S.count(k) for k in ("Np", "Nu", "Nv", "Nw", "Nqu", "Nqv", "Nqw", "n_pruned")}
    banner = [
        "=" * 60,
        "3D Schaefer-Turek channel with immersed full-height cylinder",
        "Multi-step reused-object device-solve code (CPU assembly, GPU momentum+pressure dense ops)",
        "Grid                  : 32 x 12 x 12 pressure cells",
        "Domain                : [0,2.2] x [0,0.41] x [0,0.41]",
        "Cylinder center/rad   : (0.200000, 0.200000), 0.050000",
        "nu                    : 1.000000e-03",
        "dt                    : 2.000000e-03",
        "nsteps                : 4",
        "Pressure cells        : %d" % c["Np"],
        "Active u/v/w          : %d / %d / %d" % (c["Nu"], c["Nv"], c["Nw"]),
        "Projection q faces    : %d / %d / %d" % (c["Nqu"], c["Nqv"], c["Nqw"]),
        "Pruned isolated x-y columns: %d" % c["n_pruned"],
        "Momentum solver       : explicit row-major Struct GMRES + Struct Jacobi (DEVICE mode, CPU assembly, "
        "reused objects, GPU momentum dense ops)",
        "Pressure solver       : explicit row-major Struct GMRES + Struct PFMG on sign-flipped exact A_p (DEVICE "
        "mode, CPU assembly, reused objects)",
        "Pressure matrix build : once",
        "VTK writing           : enabled (every 3 step(s), base=%s)" % base,
        "=" * 60,
        "step  ramp         max|u|         max|v|         max|w|         divL2          its(u,v,w,p)"
        "                relres(u,v,w,p)",
    ]
    assert lines[1:1 + len(banner)] == banner
    rows = _step_lines(out)
    assert [r[0] for r in rows] == [1, 2, 3, 4]  # print_every 1 (the default)
    for (step, ramp, maxes, its, rel), (s, _) in zip(rows, oracle_run):
        assert ramp == pytest.approx(s["ramp"], abs=1e-6)
        for got, k in zip(maxes[:3], ("umax", "vmax", "wmax")):
            assert got == pytest.approx(s[k], rel=2e-6, abs=1e-300), (step, k)
        assert all(abs(a - b) <= 1 for a, b in zip(its, [s["its_" + q] for q in "uvwp"])), (step, its)
        assert max(rel) <= 1e-11
    tail = lines[-7:]
    assert tail[0] == "=" * 60 and tail[1] == "Timing summary" and tail[-1] == "=" * 60
    for line, label in zip(tail[2:6], ("momentum build time    : ", "momentum solve time    : ",
                                       "pressure solve time    : ", "total wall time        : ")):
        assert line.startswith(label) and line.endswith(" s")
    # VTK files at steps 1, 3 (every 3) and 4 (the last), named with the step suffix
    names = sorted(os.listdir(tmp_path))
    assert names == ["chan_000001.vtk", "chan_000003.vtk", "chan_000004.vtk", "fields.bin"]
    r = F.read_fields(dump)
    assert r["step"] == 4
    s, fo = oracle_run[-1]
    for f in "uvwp":
        assert rel_l2(r[f], fo[f]) < 1e-9, f
    ref = str(tmp_path / "ref.vtk")
    O.write_vtk_cellcentered(ref, r["u"], r["v"], r["w"], r["p"], S.get("fluid_p"), 32, 12, 12, 2.2, 0.41, 0.41)
    assert open(ref, "rb").read() == open(str(tmp_path / "chan_000004.vtk"), "rb").read()


def test_driver_without_vtk(torch_cuda, tmp_path):
    out = _run([MAIN] + _argv(dict(FLAGS, nsteps=2)) + ["-write-vtk", "0", "-print-every", "5"], timeout=60,
               cwd=str(tmp_path))
    assert "VTK writing           : disabled" in out.splitlines()
    assert [r[0] for r in _step_lines(out)] == [1, 2]  # steps 1 and nsteps are always printed
    assert not os.listdir(tmp_path)


def test_reference_shaped_caller_matches_oracle(torch_cuda, oracle_run, tmp_path):
    dump = str(tmp_path / "caller.bin")
    out = _run([CALLER, dump] + _argv(FLAGS))
    rows = _step_lines(out)
    assert [r[0] for r in rows] == [1, 2, 3, 4]
    for (step, ramp, maxes, its, rel), (s, _) in zip(rows, oracle_run):
        assert all(abs(a - b) <= 1 for a, b in zip(its, [s["its_" + q] for q in "uvwp"])), (step, its)
        for got, k in zip(maxes[:3], ("umax", "vmax", "wmax")):
            assert got == pytest.approx(s[k], rel=2e-6, abs=1e-300), (step, k)
    r = F.read_fields(dump)
    _, fo = oracle_run[-1]
    for f in "uvwp":
        assert rel_l2(r[f], fo[f]) < 1e-9, f

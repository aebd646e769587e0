"""On-disk formats without a GPU: the cell-centred VTK file of the reference
driver (immersed_operators.hpp:1055-1076) written by libfpc's host writer is
byte for byte the file the oracle's fprintf restatement writes, on random
fields with signed zeros, subnormals, huge and tiny values; the "_%06d" step
file name (solver_helpers.hpp:470-482); the raw field dump round trip; the
output flags of the reference's command line (common.hpp:121-125)."""
import os

import numpy as np
import pytest

import fpc_amd as F
import oracle as O


def _fields(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    sizes = [(nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1), nx * ny * nz]
    out = []
    for n in sizes:
        a = rng.standard_normal(n) * 10.0 ** rng.integers(-12, 12, n)
        a[rng.choice(n, 3, replace=False)] = [0.0, -0.0, 5e-324]
        a[rng.choice(n, 2, replace=False)] = [1.7976931348623157e308 / 4, -2.2250738585072014e-308]
        out.append(a)
    fluid = (rng.random(nx * ny * nz) > 0.2).astype(np.uint8)
    return out, fluid


@pytest.mark.parametrize("dims", [(7, 5, 3), (32, 12, 12)])
def test_vtk_bytes_equal_reference_writer(tmp_path, dims):
    nx, ny, nz = dims
    (u, v, w, p), fluid = _fields(nx, ny, nz, seed=sum(dims))
    a, b = tmp_path / "lib.vtk", tmp_path / "ref.vtk"
    F.write_vtk_cellcentered(str(a), u, v, w, p, fluid, nx, ny, nz, 2.5, 0.41, 0.41)
    O.write_vtk_cellcentered(str(b), u, v, w, p, fluid, nx, ny, nz, 2.5, 0.41, 0.41)
    ba, bb = a.read_bytes(), b.read_bytes()
    assert len(ba) == len(bb)
    assert ba == bb
    lines = ba.decode().splitlines()
    assert lines[0] == "# vtk DataFile Version 3.0"
    assert lines[4] == "DIMENSIONS %d %d %d" % (nx + 1, ny + 1, nz + 1)
    assert len(lines) == 8 + 5 * 2 + 5 * nx * ny * nz


def test_vtk_step_file_names():
    cases = [("st3d_channel_cylinder_struct_mac_cuda.vtk", 1), ("out.vtk", 123456), ("out.vtk", 1234567),
             ("noext", 7), ("a.b/c", 3), ("dir/run.1.vtk", 42), (".vtk", 0)]
    for base, step in cases:
        assert F.vtk_filename_with_step(base, step) == O.vtk_filename_with_step(base, step), (base, step)
    assert F.vtk_filename_with_step("x.vtk", 5) == "x_000005.vtk"


def test_field_dump_round_trip(tmp_path):
    nx, ny, nz = 9, 4, 3
    (u, v, w, p), _ = _fields(nx, ny, nz, seed=3)
    path = str(tmp_path / "f.bin")
    F.write_fields(path, 17, nx, ny, nz, u, v, w, p)
    r = F.read_fields(path)
    assert (r["nx"], r["ny"], r["nz"], r["step"]) == (nx, ny, nz, 17)
    for name, a in zip("uvwp", (u, v, w, p)):
        assert np.array_equal(r[name].view(np.uint64), a.view(np.uint64)), name
    assert os.path.getsize(path) == 40 + 8 * (len(u) + len(v) + len(w) + len(p))


def test_output_flags_parsed():
    p = F.default_params()
    assert (p.write_vtk, p.vtk_every, p.device) == (1, 0, 0)  # common.hpp:60, 82-83
    assert p.vtk_filename == b"st3d_channel_cylinder_struct_mac_cuda.vtk"
    p = F.parse_args("-write-vtk 0 -vtk-every 5 -vtk run/out.vtk -device 0 -nx 16".split())
    assert (p.write_vtk, p.vtk_every, p.device, p.nx) == (0, 5, 0, 16)
    assert p.vtk_filename == b"run/out.vtk"

"""Writes tests/golden/cfg2_3steps.npz: the oracle's state after the first three
time steps of the production channel 512x96x96 (Um 2.25, dt 2.5e-4, GMRES(30),
inlet profile plus seeded noise 1e-3*Um, seed 0), solved at rtol 1e-10 (the
parameters are tests/cfg2_case.py's CFG2_PARITY).

Stored per field f in u, v, w, p:
  f_zmid    the plane k = nz/2 over the x window [X0, X1) (faces / cells), all y
  f_xcyl    the plane through the cylinder axis (face / cell i = XC), all y, z
  f_knorm   the L2 norm of every z-plane of the full field (a checksum of all planes)
  f_norm    the L2 norm of the full field
plus cd, cl and the per-step iteration counts of the four solves.

The oracle runs with OpenMP over its solver and stage loops (set_threads);
its results depend on the thread count only at the rounding level (fixed
per-thread partial sums), far below the bar.
Run: python tests/golden/make_cfg2_fixture.py [threads]   (~10 min, ~25 GB of host memory)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
sys.path.insert(0, os.path.join(HERE, ".."))
import oracle as O  # noqa: E402
from cfg2_case import CFG2_PARITY, NSTEPS, field_views  # noqa: E402


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else os.cpu_count()
    O.set_threads(threads)
    t0 = time.time()
    sim = O.Sim(O.default_params(**CFG2_PARITY))
    print(f"setup {time.time() - t0:.1f} s", flush=True)
    its = []
    for n in range(NSTEPS):
        t1 = time.time()
        sim.step(1)
        s = sim.scalars()
        its.append([s["its_u"], s["its_v"], s["its_w"], s["its_p"]])
        print(f"step {n + 1}: its {its[-1]} relres_p {s['relres_p']:.3e} cd {s['cd']:.12g} "
              f"({time.time() - t1:.1f} s)", flush=True)
    out = {}
    p = CFG2_PARITY
    for f in ("u", "v", "w", "p"):
        out.update(field_views(f, sim.get(f), p["nx"], p["ny"], p["nz"]))
    s = sim.scalars()
    out["cd"] = np.float64(s["cd"])
    out["cl"] = np.float64(s["cl"])
    out["its"] = np.array(its, dtype=np.int32)
    out["threads"] = np.int32(threads)
    np.savez_compressed(os.path.join(HERE, "cfg2_3steps.npz"), **out)
    print("wrote cfg2_3steps.npz", {k: v.shape for k, v in out.items()}, f"total {time.time() - t0:.1f} s")


if __name__ == "__main__":
    main()

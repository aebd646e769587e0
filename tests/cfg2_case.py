"""The production channel 512x96x96 as a parity case: parameters of its first
steps at a tight solver tolerance, and the views of a field kept in the
committed fixture (tests/golden/cfg2_3steps.npz)."""
import numpy as np

# 512x96x96, Um 2.25 (Re ~100), dt 2.5e-4, GMRES(30); inlet profile + uniform noise of
# amplitude 1e-3 * Um, seed 0.  rtol 1e-10 instead of the production 1e-8.  Not 1e-12:
# there the first pressure solves (the projection of the noisy initial field) stall at the
# attainable accuracy, a true relative residual of ~2e-12 after a restart, and the
# iteration count at which a run gives up depends on rounding (40 against 45 its).  At
# 1e-10 the solves end well above that floor, with the same iteration counts as the oracle.
CFG2_PARITY = dict(nx=512, ny=96, nz=96, Lx=2.5, Ly=0.41, Lz=0.41, xcirc=0.5, ycirc=0.2, Rcirc=0.05, Um=2.25,
                   nu=1e-3, rho=1.0, dt=2.5e-4, vel_kdim=30, p_kdim=30, vel_tol=1e-10, p_tol=1e-10, vel_maxit=500,
                   p_maxit=500, init_profile=1, noise_amp=1e-3 * 2.25, seed=0)
NSTEPS = 3
# x window of the z-mid plane (the cylinder at x = 0.5 sits at i ~ 102) and the x index
# of the plane through the cylinder axis
X0, X1, XC = 64, 192, 102

# staggered extents (nx', ny', nz') of each field
def extents(f, nx, ny, nz):
    return {"u": (nx + 1, ny, nz), "v": (nx, ny + 1, nz), "w": (nx, ny, nz + 1), "p": (nx, ny, nz)}[f]


def field_views(f, a, nx, ny, nz):
    """The fixture's views of field f (flat, x fastest)."""
    ex, ey, ez = extents(f, nx, ny, nz)
    a3 = np.asarray(a, dtype=np.float64).reshape(ez, ey, ex)
    return {
        f + "_zmid": np.ascontiguousarray(a3[ez // 2, :, X0:X1]),
        f + "_xcyl": np.ascontiguousarray(a3[:, :, XC]),
        f + "_knorm": np.sqrt(np.einsum("kji,kji->k", a3, a3)),
        f + "_norm": np.float64(np.linalg.norm(a3)),
    }

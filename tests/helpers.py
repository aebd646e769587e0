"""Test helpers: parameter sets of the benchmark configurations, Struct
matrix conversion for scipy checks, and small comparisons."""
import numpy as np
import scipy.sparse as sp

OFFS = [(0, 0, 0), (-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]

# Configuration 1 of the benchmark (reference CPU case), 128x32x32.
CFG1 = dict(nx=128, ny=32, nz=32, Lx=2.5, Ly=0.41, Lz=0.41, xcirc=0.5, ycirc=0.2, Rcirc=0.05, Um=0.45, nu=1e-3,
            rho=1.0, dt=1e-3, vel_tol=1e-8, p_tol=1e-8)
# Small variants used by fast tests.
TINY = dict(nx=32, ny=12, nz=12)
SMALL = dict(nx=48, ny=16, nz=12, Lx=1.0, xcirc=0.3, ycirc=0.2, Rcirc=0.07, Um=0.45, dt=2e-3, vel_tol=1e-10,
             p_tol=1e-10, vel_kdim=30, p_kdim=30)


def box_csr(nx, ny, nz, vals):
    """scipy CSR of a 7-point box matrix given in AoS (7 per point); out-of-box couplings dropped."""
    N = nx * ny * nz
    a = np.asarray(vals).reshape(N, 7)
    idx = np.arange(N).reshape(nz, ny, nx)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    rows, cols, data = [], [], []
    for s, (di, dj, dk) in enumerate(OFFS):
        ii, jj, kk = i + di, j + dj, k + dk
        m = (ii >= 0) & (ii < nx) & (jj >= 0) & (jj < ny) & (kk >= 0) & (kk < nz)
        r = idx[m]
        rows.append(r)
        cols.append(idx[kk[m], jj[m], ii[m]])
        data.append(a[r, s])
    return sp.csr_matrix((np.concatenate(data), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))


def csr_1based(rowptr, colind, vals, ncols):
    """scipy CSR of a 1-based-column matrix (reference ManagedCSR)."""
    m = sp.csr_matrix((vals, np.asarray(colind) - 1, rowptr), shape=(len(rowptr) - 1, ncols))
    return m


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))

"""Numpy restatement of the flow-file conversions of the reference, for one rank,
and of poisson_p_chan (the pressure ReadHDF5 recomputes for a file without P).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of libdgpu_io.so.

Arrays are the reference's physical layout (0:Nx+1, 0:Nzp+1, 0:Nyp+1) as C-order
numpy arrays of shape (Ny+2, Nz+2, Nx+2), i.e. a[y, z, x].
"""
from __future__ import annotations

import numpy as np


def swapzy(a: np.ndarray, nx: int, ny: int, nz: int) -> np.ndarray:
    """for/tools.f90:216-232: out(x+1, y, z+1) = in(x, z, y) for x < Nx, y = 1..Nyp,
    z < Nzp.  Returned as the C-order file array [z, y-1, x]."""
    return np.ascontiguousarray(a[1:ny + 1, :nz, :nx].transpose(1, 0, 2))


def swapyz(f: np.ndarray, nx: int, ny: int, nz: int) -> np.ndarray:
    """for/tools.f90:235-251: the physical array zeroed, rows 1..Nyp of x < Nx,
    z < Nzp taken from the file array."""
    out = np.zeros((ny + 2, nz + 2, nx + 2))
    out[1:ny + 1, :nz, :nx] = f.transpose(1, 0, 2)
    return out


def g2gf(a: np.ndarray, nx: int, ny: int, nz: int) -> np.ndarray:
    """for/tools.f90:299-315 at one rank (rankY = 0 = NprocY - 1): the wall rows
    copied in from rows 2 and Nyp, then each row 1..Nyp replaced by its mean with
    the row above, in increasing y (each mean uses the row above as it still was)."""
    v = a.copy()
    s = (slice(None), slice(0, nz), slice(0, nx))
    v[1][s[1:]] = v[2][s[1:]]
    v[ny + 1][s[1:]] = v[ny][s[1:]]
    # increasing y: row y reads row y + 1 before it is updated, so a vector form holds
    v[1:ny + 1, :nz, :nx] = 0.5 * (v[1:ny + 1, :nz, :nx] + v[2:ny + 2, :nz, :nx])
    return v


def gf2g(a: np.ndarray, nx: int, ny: int, nz: int) -> np.ndarray:
    """for/tools.f90:255-295 at one rank: row Nyp+1 = row Nyp, then for y = Nyp
    down to 1, row y = 2 row y - row y+1 (a recurrence on the updated row above);
    finally rows 1 and Nyp+1 of the whole array set to minus rows 2 and Nyp."""
    v = a.copy()
    v[ny + 1, :nz, :nx] = v[ny, :nz, :nx]
    for y in range(ny, 0, -1):
        v[y, :nz, :nx] = 2.0 * v[y, :nz, :nx] - v[y + 1, :nz, :nx]
    v[1] = -v[2]
    v[ny + 1] = -v[ny]
    return v


def poisson_source(o):
    """cs1 of poisson_p_chan (for/solver.f90:1390-1530) for the oracle's state, the
    reference's way: spectral derivatives, products in physical space, each
    product transformed back; rows 2..Nyp, then the scalars' y gradient (except the
    (0, 0) mode).  Compact Fourier arrays [j][k][i]."""
    p, cfg = o.p, o.p.cfg
    s = o.download_state()
    ny = cfg.ny
    dyf, gyf = p.met["dyf"], p.met["gyf"]
    kx = p.kx[:cfg.kxr][None, :]
    kz = p.kz[:, None]
    J = np.arange(2, ny + 1)
    phys, four = o.fft_to_physical, o.fft_to_fourier

    def field(rows_fn):
        a = np.zeros_like(s.cu1)
        for j in J:
            a[j] = rows_fn(j)
        return phys(a)

    def rowsonly(f):
        g = np.zeros_like(f)
        g[J] = f[J]
        return g
    u1, u2, u3 = s.cu1, s.cu2, s.cu3
    f1 = field(lambda j: 1j * kx * u1[j])
    f2 = field(lambda j: (u2[j + 1] - u2[j]) / dyf[j])
    f3 = field(lambda j: 1j * kz * u3[j])
    cs1 = four(rowsonly(f1 ** 2)) + four(rowsonly(f2 ** 2)) + four(rowsonly(f3 ** 2))
    for a_fn, b_fn in ((lambda j: (u1[j + 1] - u1[j - 1]) / (2.0 * dyf[j]), lambda j: 1j * kx * 0.5 * (u2[j] + u2[j + 1])),
                       (lambda j: (u3[j + 1] - u3[j - 1]) / (2.0 * dyf[j]), lambda j: 1j * kz * 0.5 * (u2[j] + u2[j + 1])),
                       (lambda j: 1j * kz * u1[j], lambda j: 1j * kx * u3[j])):
        fa, fb = field(a_fn), field(b_fn)
        cs1 = cs1 + four(rowsonly(2.0 * fa * fb))
    out = np.zeros_like(cs1)
    out[J] = cs1[J]
    for t in s.cth:
        g = (t[J + 1] - t[J - 1]) / (gyf[J + 1] - gyf[J - 1])[:, None, None]
        g[:, 0, 0] = 0.0
        out[J] = out[J] + g
    return out


def poisson_p(o):
    """cp of poisson_p_chan (for/solver.f90:1532-1560, BC rows of
    apply_BC_poisson_mpi 2375-2411): per (kx, kz) column the tridiagonal system on
    rows 1..Nyp solved by the Thomas algorithm; returns the compact cp with rows
    1..Nyp replaced."""
    p, cfg = o.p, o.p.cfg
    ny, n2 = cfg.ny, cfg.ny2
    dy, dyf = p.met["dy"], p.met["dyf"]
    cs1 = poisson_source(o)
    kx2 = (p.kx[:cfg.kxr] ** 2)[None, :]
    kz2 = (p.kz ** 2)[:, None]
    shp = cs1.shape[1:]
    # rows 0..Nyp+1 start as (0, 1, 0 | 0); rows 2..Nyp from the Laplacian
    l = np.zeros((n2,) + shp); d = np.ones((n2,) + shp); u = np.zeros((n2,) + shp)
    v = np.zeros((n2,) + shp, np.complex128)
    for j in range(2, ny + 1):
        l[j] = 1.0 / (dy[j] * dyf[j])
        d[j] = -kx2 - kz2 - 1.0 / (dy[j + 1] * dyf[j]) - 1.0 / (dy[j] * dyf[j])
        u[j] = 1.0 / (dy[j + 1] * dyf[j])
        v[j] = -1.0 * cs1[j]
    # apply_BC_poisson_mpi: row 1 Neumann (Dirichlet in the (0, 0) mode), row Nyp Neumann
    d[1] = 1.0; u[1] = -1.0; u[1][0, 0] = 0.0; v[1] = 0.0
    d[ny] = -1.0; l[ny] = 1.0; v[ny] = 0.0
    # thomas_forward/backward_complex_mpi at one rank: rows 1..Nyp (u of row Nyp unused)
    for j in range(2, ny + 1):
        m = -l[j] / d[j - 1]
        d[j] = d[j] + m * u[j - 1]
        v[j] = v[j] + m * v[j - 1]
    v[ny] = v[ny] / d[ny]
    for j in range(ny - 1, 0, -1):
        v[j] = (v[j] - u[j] * v[j + 1]) / d[j]
    cp = o.download_state().cp.copy()
    cp[1:ny + 1] = v[1:ny + 1]
    return cp

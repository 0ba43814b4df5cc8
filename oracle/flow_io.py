"""Numpy restatement of the flow-file conversions of the reference, for one rank.

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

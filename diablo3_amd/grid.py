"""Wall-normal grid, grid metrics and wavenumbers.

* :func:`create_grid` restates proc/init/create_grid.m (the MATLAB script that writes
  ``grid.h5:/grids/y``) for every GRID_TYPE; the configs use GRID_TYPE 1.
* :func:`metrics` restates ReadGridHDF5 (for/domain.f90:533-564) and
  create_grid_chan (for/domain.f90:276-300) at one rank.
* :func:`wavenumbers` restates init_fft (for/fft.f90:147, 209-225), including its
  single-precision pi.
"""
from __future__ import annotations

import math

import numpy as np


def create_grid(n: int, length: float, cs: float, grid_type: int = 1) -> np.ndarray:
    """Return the n+1 values G(2..N+2) the script writes to grid.h5:/grids/y."""
    G = np.zeros(n + 3)
    GF = np.zeros(n + 3)
    for J in range(1, n + 2):
        if grid_type == 1:
            G[J + 1] = (length / 2.0) * math.tanh(cs * ((2.0 * (J - 1)) / n - 1.0)) / math.tanh(cs)
        elif grid_type == 2:
            G[J + 1] = (length / 2.0) * (cs * ((2.0 * (J - 1)) / n - 1.0) ** 3
                                         + (2.0 * (J - 1)) / n - 1.0) / (cs + 1)
        elif grid_type == 3:
            G[J + 1] = (length / 2.0) * math.tanh(cs * ((2.0 * (J - 1)) / n - 1.0)) / math.tanh(cs) \
                + length / 2.0
        elif grid_type == 4:
            G[J + 1] = length * math.tanh(cs * ((n - (J - 1)) / n - 1.0)) / math.tanh(cs) + length
        else:
            raise ValueError("unknown GRID_TYPE")
    if grid_type == 4:
        G = -G
    for J in range(1, n + 1):
        GF[J + 1] = (G[J + 1] + G[J + 2]) / 2.0
    gf_lower, gf_upper = GF[2], GF[n + 1]
    g_lower, g_upper = G[2], G[n + 2]
    scale = (g_upper - g_lower) / (gf_upper - gf_lower)
    for J in range(1, n + 1):
        GF[J + 1] = GF[J + 1] * scale
    for J in range(1, n + 2):
        G[J + 1] = G[J + 1] * scale
    shift = g_lower - GF[2]
    for J in range(1, n + 1):
        GF[J + 1] = GF[J + 1] + shift
    for J in range(1, n + 2):
        G[J + 1] = G[J + 1] + shift
    return G[2:n + 3].copy()


def metrics(gy_file: np.ndarray) -> dict:
    """gy, gyf, dy, dyf as (0:Nyp+1) arrays from the ny+1 grid-file values.

    gy(0) is never read by the reference and stays 0 (module storage)."""
    ny = len(gy_file) - 1
    n2 = ny + 2
    gy = np.zeros(n2)
    gyf = np.zeros(n2)
    dy = np.zeros(n2)
    dyf = np.zeros(n2)
    gy[1:ny + 2] = gy_file
    for j in range(1, ny + 1):
        gyf[j] = 0.5 * (gy[j] + gy[j + 1])
    gyf[0] = 2.0 * gyf[1] - gyf[2]
    gyf[ny + 1] = 2.0 * gyf[ny] - gyf[ny - 1]
    for j in range(1, ny + 2):
        dy[j] = gyf[j] - gyf[j - 1]
    for j in range(1, ny + 1):
        dyf[j] = gy[j + 1] - gy[j]
    dyf[0] = dyf[1]
    dyf[ny + 1] = dyf[ny]
    return {"gy": gy, "gyf": gyf, "dy": dy, "dyf": dyf}


PI_F32 = float(np.float32(4.0) * np.arctan(np.float32(1.0)))


def wavenumbers(nx: int, nz: int, lx: float, lz: float):
    """kx(0:Nxp-1), kz(0:2Nkz) exactly as init_fft (pi = 4*atan(1.0) in real(4))."""
    nxp, nkz = nx // 2, nz // 3
    two_pi = 2.0 * PI_F32
    kx = np.array([(float(i) * two_pi) / lx for i in range(nxp)])
    kz = np.zeros(2 * nkz + 1)
    for k in range(nkz + 1):
        kz[k] = (float(k) * two_pi) / lz
    for k in range(1, nkz + 1):
        kz[2 * nkz + 1 - k] = -((float(k) * two_pi) / lz)
    return kx, kz


def gx(nx: int, lx: float) -> np.ndarray:
    """gx(i) = (i*Lx)/Nx (for/domain.f90:254-255)."""
    return np.array([(i * lx) / nx for i in range(nx + 2)])

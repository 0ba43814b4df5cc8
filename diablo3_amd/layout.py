"""Conversions between the Fortran declared layouts and compact numpy views.

Fourier arrays cross the C ABI as ``complex(8) cu(0:Nxp, 0:Nz+1, 0:Nyp+1)``
(for/fft.f90:71-96); in numpy (C order) that is shape ``(Ny+2, Nz+2, Nxp+1)``.
The compact view used by the host code is ``[j][k][i]`` with ``i = 0..Nkx`` and
``k = 0..2Nkz`` (the only modes the reference keeps non-zero).

Physical arrays are ``real(8) u(0:Nx+1, 0:Nzp+1, 0:Nyp+1)``: numpy shape
``(Ny+2, Nz+2, Nx+2)``; compact view ``[j][z][x]`` with x < Nx, z < Nz.
"""
from __future__ import annotations

import numpy as np

from .config import Config


def fortran_fourier_shape(cfg: Config):
    return (cfg.ny2, cfg.nz + 2, cfg.nxp + 1)


def fortran_physical_shape(cfg: Config):
    return (cfg.ny2, cfg.nz + 2, cfg.nx + 2)


def to_fortran_fourier(cfg: Config, compact: np.ndarray) -> np.ndarray:
    out = np.zeros(fortran_fourier_shape(cfg), dtype=np.complex128)
    out[:, :cfg.kzr, :cfg.kxr] = compact
    return out


def from_fortran_fourier(cfg: Config, fort: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(fort[:, :cfg.kzr, :cfg.kxr])


def to_fortran_physical(cfg: Config, compact: np.ndarray) -> np.ndarray:
    out = np.zeros(fortran_physical_shape(cfg), dtype=np.float64)
    out[:, :cfg.nz, :cfg.nx] = compact
    return out


def from_fortran_physical(cfg: Config, fort: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(fort[:, :cfg.nz, :cfg.nx])


def compact_zeros(cfg: Config) -> np.ndarray:
    return np.zeros((cfg.ny2, cfg.kzr, cfg.kxr), dtype=np.complex128)


def fft_xz_to_fourier(cfg: Config, phys: np.ndarray) -> np.ndarray:
    """Host restatement of fft_xz_to_fourier (for/fft.f90:242-307) for the IC hooks.

    phys is [j][z][x]; returns the compact [j][k][i] spectrum: x r2c / Nx with
    i > Nkx zeroed, z forward / Nz, negative kz packed into k = Nkz+1..2Nkz."""
    nx, nz, nkx, nkz = cfg.nx, cfg.nz, cfg.nkx, cfg.nkz
    h = np.fft.rfft(phys, axis=2)[:, :, :nkx + 1] / nx
    f = np.fft.fft(h, axis=1)
    out = np.empty((phys.shape[0], cfg.kzr, nkx + 1), dtype=np.complex128)
    out[:, :nkz + 1] = f[:, :nkz + 1] / nz
    out[:, nkz + 1:] = f[:, nz - nkz:] / nz
    return out


def fft_xz_to_physical(cfg: Config, four: np.ndarray) -> np.ndarray:
    """Host restatement of fft_xz_to_physical (for/fft.f90:311-374)."""
    nx, nz, nkx, nkz = cfg.nx, cfg.nz, cfg.nkx, cfg.nkz
    full = np.zeros((four.shape[0], nz, nkx + 1), dtype=np.complex128)
    full[:, :nkz + 1] = four[:, :nkz + 1]
    full[:, nz - nkz:] = four[:, nkz + 1:]
    h = np.fft.ifft(full, axis=1) * nz
    hx = np.zeros((four.shape[0], nz, nx // 2 + 1), dtype=np.complex128)
    hx[:, :, :nkx + 1] = h
    hx[:, :, 0] = hx[:, :, 0].real
    return np.fft.irfft(hx, n=nx, axis=2) * nx

"""Initial-condition hooks: create_flow_chan / create_th_chan, kept host-side.

The reference builds the initial state on the host in physical space, transforms it,
applies the Fourier-space boundary posts and removes the divergence
(for/flow.f90:260-499, 562-659).  These functions restate that sequence up to the
projection, which the caller runs on the device (``Solver.rem_div``), exactly where
create_flow_chan calls rem_div_chan (for/flow.f90:487).

Differences, stated once:
  * the reference draws its noise from the compiler's ``random_number`` stream
    seeded with ``rank*k+i+999`` (for/flow.f90:265-274), which no other toolchain
    reproduces; here the same draws, in the same loop order, come from numpy's
    PCG64 seeded by the caller;
  * ``p`` is transformed uninitialised by the reference (for/flow.f90:433, 441);
    here cp starts at zero;
  * planes the reference leaves unset (e.g. th at j = Nyp+1 for IC type 4) are zero.
  * ``theta_noise`` adds seeded uniform noise to the scalar on rows 1..Nyp, which the
    reference never does; it is off unless a config asks for it.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List

import numpy as np

from .config import DIRICHLET, NEUMANN, Config
from .grid import gx as grid_gx
from .layout import fft_xz_to_fourier


@dataclass
class FourierState:
    cu1: np.ndarray
    cu2: np.ndarray
    cu3: np.ndarray
    cp: np.ndarray
    cth: List[np.ndarray]

    def copy(self) -> "FourierState":
        return FourierState(self.cu1.copy(), self.cu2.copy(), self.cu3.copy(),
                            self.cp.copy(), [t.copy() for t in self.cth])


def create_th_chan(cfg: Config, gyf: np.ndarray) -> List[np.ndarray]:
    """Physical scalar fields [j][z][x] (for/flow.f90:562-659)."""
    ny, nz, nx, ly, lx = cfg.ny, cfg.nz, cfg.nx, cfg.ly, cfg.lx
    gxv = grid_gx(nx, lx)[:nx]
    out = []
    pi = 4.0 * math.atan(1.0)
    for n in range(cfg.n_th):
        th = np.zeros((cfg.ny2, nz, nx))
        ri = cfg.ri[n]
        if not cfg.create_new_th[n]:
            out.append(th)
            continue
        t = cfg.ic_type
        for j in range(1, ny + 1):
            if t == 0:
                th[j] = (ri * np.sin(2.0 * pi * gxv / lx) / (4.0 * pi ** 2))[None, :]
            elif t in (1, 2):
                if cfg.th_bc_ymin[n] == DIRICHLET and cfg.th_bc_ymax[n] == DIRICHLET:
                    val = ((cfg.th_bc_ymax_c1[n] - cfg.th_bc_ymin_c1[n]) * (gyf[j] + 1.0) / 2.0
                           + cfg.th_bc_ymin_c1[n]) if gyf[j] <= 2.0 else cfg.th_bc_ymax_c1[n]
                    th[j] = val
                elif cfg.th_bc_ymin[n] == NEUMANN and cfg.th_bc_ymax[n] == NEUMANN:
                    th[j] = cfg.th_bc_ymin_c1[n] * gyf[j]
            elif t == 3:
                th[j] = ri * math.tanh(gyf[j])
            elif t in (4, 5):
                th[j] = ri * (gyf[j] - 0.5 * ly)
            elif t in (6, 7, 8):
                if ri == 0:
                    th[j] = (cfg.ro_inv * np.tanh((gxv - 0.5 * lx) / cfg.delta)
                             / math.tanh(0.5 * lx / cfg.delta)
                             + cfg.ro_inv * (1.0 - 2.0 / lx * gxv))[None, :]
                else:
                    s = math.sqrt(ri ** 2 + cfg.ro_inv ** 2)
                    th[j] = (s * np.tanh(((gxv - 0.5 * lx) * cfg.ro_inv + (gyf[j] - 0.5 * ly) * ri)
                                         / (cfg.delta * s))
                             + cfg.ro_inv * (1.0 - 2.0 / lx * gxv))[None, :]
            else:
                raise ValueError("unsupported IC_type in create_th")
        out.append(th)
    return out


def create_flow_physical(cfg: Config, gyf: np.ndarray, rng: np.random.Generator):
    """Physical u1, u2, u3 [j][z][x] for IC types 0-8 plus physical noise
    (for/flow.f90:283-426)."""
    ny, nz, nx, ly, lx, dl = cfg.ny, cfg.nz, cfg.nx, cfg.ly, cfg.lx, cfg.delta
    gxv = grid_gx(nx, lx)[:nx]
    sh = (cfg.ny2, nz, nx)
    u1, u2, u3 = np.zeros(sh), np.zeros(sh), np.zeros(sh)
    t = cfg.ic_type
    if t == 0:
        for j in range(0, ny + 1):
            u1[j] = (3.0 / 2.0) * cfg.ubulk0 * (1.0 - gyf[j] ** 2)
    elif t == 1:
        for j in range(1, ny + 1):
            u1[j] = -(3.0 / 2.0) * cfg.ubulk0 * gyf[j] ** 2 + 3.0 * cfg.ubulk0 * gyf[j]
    elif t == 2:
        for j in range(0, ny + 1):
            u1[j] = gyf[j]
    elif t == 3:
        for j in range(0, ny + 1):
            u1[j] = math.tanh(gyf[j])
    elif t == 4:
        pass
    elif t == 5:
        a = (math.tanh(0.5 * lx / dl) - math.tanh(-0.5 * lx / dl)) / lx
        prof = (np.tanh((gxv - 0.5 * lx) / dl) - a * (gxv - 0.5 * lx)) / (dl * (1.0 / dl - a))
        for j in range(0, ny + 1):
            u3[j] = prof[None, :]
    elif t == 6:
        ri = cfg.ri[0]
        for j in range(0, ny + 1):
            if ri == 0:
                u3[j] = ((gyf[j] - 0.5 * ly) / np.cosh((gxv - 0.5 * lx) / dl) ** 2
                         / math.tanh(0.5 * lx / dl) - 2.0 * dl / lx * (gyf[j] - 0.5 * ly))[None, :]
            else:
                s = math.sqrt(ri ** 2 + cfg.ro_inv ** 2)
                u3[j] = (dl * s / ri * np.tanh(((gxv - 0.5 * lx) * cfg.ro_inv + (gyf[j] - 0.5 * ly) * ri)
                                               / (dl * s))
                         - 2.0 * dl / lx * (gyf[j] - 0.5 * ly))[None, :]
    elif t == 7:
        for j in range(0, ny + 1):
            u3[j] = (gyf[j] / np.cosh((gxv - 0.5 * lx) / dl) ** 2 / math.tanh(0.5 * lx / dl)
                     - 2.0 * dl / lx * gyf[j])[None, :]
    elif t == 8:
        for j in range(0, ny + 1):
            u3[j] = -2.0 * dl / lx * (gyf[j] - 0.5 * ly)
    else:
        raise ValueError("unsupported IC_type in create_flow")
    if cfg.physical_noise:
        rng.random(3)  # the three discarded draws of for/flow.f90:409-411
        # draws in the loop order j, z, x, (u1, u2, u3), one plane at a time
        for j in range(cfg.ny2):
            r = rng.random((nz, nx, 3))
            u1[j] += cfg.kick * (r[..., 0] - 0.5)
            u2[j] += cfg.kick * (r[..., 1] - 0.5)
            u3[j] += cfg.kick * (r[..., 2] - 0.5)
    return u1, u2, u3


def apply_bc_vel_post(cfg: Config, met: dict, s: FourierState, time: float = 0.0) -> None:
    """apply_BC_vel_lower_post / apply_BC_vel_upper_post (for/flow.f90:1161-1360)
    on compact Fourier arrays, one rank (rankZ = 0 owns the mean mode)."""
    ny, dy, dyf = cfg.ny, met["dy"], met["dyf"]
    cu1, cu2, cu3 = s.cu1, s.cu2, s.cu3
    # lower
    if cfg.u_bc_ymin == DIRICHLET:
        cu1[0] = 0.0
        cu1[1] = 0.0
        cu1[1, 0, 0] = cfg.u_bc_ymin_c1
        cu1[0, 0, 0] = cfg.u_bc_ymin_c1
    else:
        cu1[1] = cu1[2]
        cu1[0] = cu1[1]
        cu1[1, 0, 0] = cu1[2, 0, 0] - dy[2] * cfg.u_bc_ymin_c1
        cu1[0, 0, 0] = cu1[1, 0, 0] - dy[1] * cfg.u_bc_ymin_c1
    if cfg.w_bc_ymin == DIRICHLET:
        cu3[0] = 0.0
        cu3[1] = 0.0
        cu3[1, 0, 0] = cfg.w_bc_ymin_c1
        cu3[0, 0, 0] = cfg.w_bc_ymin_c1
    else:
        cu3[1] = cu3[2]
        cu3[0] = cu3[1]
        cu3[1, 0, 0] = cu3[2, 0, 0] - dy[2] * cfg.w_bc_ymin_c1
        cu3[0, 0, 0] = cu3[1, 0, 0] - dy[1] * cfg.w_bc_ymin_c1
    if cfg.v_bc_ymin == DIRICHLET:
        cu2[1] = 2.0 * cfg.v_bc_ymin_c1 - cu2[2]
        cu2[0] = cu2[1]
    else:
        cu2[1] = cu2[2]
        cu2[0] = cu2[1]
        cu2[1, 0, 0] = cu2[2, 0, 0] - dyf[1] * cfg.v_bc_ymin_c1
        cu2[0, 0, 0] = cu2[1, 0, 0] - dyf[1] * cfg.v_bc_ymin_c1
    # upper
    if cfg.u_bc_ymax == DIRICHLET:
        cu1[ny] = 0.0
        cu1[ny + 1] = 0.0
        cu1[ny, 0, 0] = cfg.u_bc_ymax_c1
        cu1[ny + 1, 0, 0] = cfg.u_bc_ymax_c1
    else:
        cu1[ny] = cu1[ny - 1]
        cu1[ny + 1] = cu1[ny]
        cu1[ny, 0, 0] = cu1[ny - 1, 0, 0] + dy[ny] * cfg.u_bc_ymax_c1
        cu1[ny + 1, 0, 0] = cu1[ny, 0, 0] + dy[ny] * cfg.u_bc_ymax_c1
    if cfg.w_bc_ymax == DIRICHLET:
        cu3[ny] = 0.0
        cu3[ny + 1] = 0.0
        cu3[ny, 0, 0] = cfg.w_bc_ymax_c1
        cu3[ny + 1, 0, 0] = cfg.w_bc_ymax_c1
        cu3[ny + 1, 0, 0] = 0.0
    else:
        c1 = cfg.w_bc_ymax_c1
        if cfg.f_type == 4:
            c1 = cfg.w_bc_ymax_c1 + cfg.amp_omega0 * math.sin(
                cfg.omega0 * (cfg.ro_inv / cfg.delta * time - cfg.force_start))
        cu3[ny] = cu3[ny - 1]
        cu3[ny + 1] = cu3[ny]
        cu3[ny, 0, 0] = cu3[ny - 1, 0, 0] + dy[ny] * c1
        cu3[ny + 1, 0, 0] = cu3[ny, 0, 0] + dy[ny] * c1
    if cfg.v_bc_ymax == DIRICHLET:
        # the reference sets only the mean mode here (for/flow.f90:1338-1344)
        cu2[ny + 1, 0, 0] = 2.0 * cfg.v_bc_ymax_c1 - cu2[ny, 0, 0]
    else:
        cu2[ny] = cu2[ny - 1]
        cu2[ny + 1] = cu2[ny]
        cu2[ny + 1, 0, 0] = cu2[ny, 0, 0] + dy[ny] * cfg.v_bc_ymax_c1


def initial_state(cfg: Config, met: dict, seed: int, theta_noise: float = 0.0) -> FourierState:
    """create_th_chan + create_flow_chan up to (not including) rem_div_chan.

    Returns the compact Fourier state.  The caller projects it (Solver.rem_div)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    gyf = met["gyf"]
    ths = create_th_chan(cfg, gyf)
    if theta_noise:
        for th in ths:
            th[1:cfg.ny + 1] += theta_noise * (2.0 * rng.random(th[1:cfg.ny + 1].shape) - 1.0)
    cth = [fft_xz_to_fourier(cfg, th) for th in ths]
    u1, u2, u3 = create_flow_physical(cfg, gyf, rng)
    s = FourierState(fft_xz_to_fourier(cfg, u1), fft_xz_to_fourier(cfg, u2),
                     fft_xz_to_fourier(cfg, u3),
                     np.zeros((cfg.ny2, cfg.kzr, cfg.kxr), dtype=np.complex128), cth)
    apply_bc_vel_post(cfg, met, s)
    return s

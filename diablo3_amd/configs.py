"""The named workloads, as :class:`Config` factories.

Code axes throughout: (x, y, z) with y wall-normal; y in [-1, 1] is GRID_TYPE 1 of
proc/init/create_grid.m with Lz(user) = 2 and stretch CS = 1.75.

Front physics, written in terms of the reference's knobs:
  * Coriolis f = Ro_inv / delta (for/solver.f90:502-509); delta = 1, Ro_inv = f.
  * Background buoyancy gradient M^2 in x = dTHdX (for/solver.f90:595-602); the
    thermal wind V_g(y) = dTHdX (y - Ly/2) delta / Ro_inv along z is applied by
    user_rhs_chan_fourier (for/solver.f90:47-102).  Infinite-front IC type 4 would
    tie dTHdX to f (for/parameters.f90:257-265); these configs set it explicitly.
  * N^2: the scalar starts as theta = Ri(1) * (y - Ly/2) (IC type 4,
    for/flow.f90:620-628) with Ri(1) = N^2, and gravity grav_y = 1, so the
    buoyancy term grav_y*theta (for/solver.f90:518-538) has N^2 as its gradient.
  * Bottom no-slip: u1, u2, u3 Dirichlet 0.  Top free-slip: u1 Neumann 0, u2
    Dirichlet 0, u3 Neumann with the thermal-wind compensation -M^2 delta/Ro_inv
    (the form of for/parameters.f90:259-263).  Scalar: Neumann dtheta/dy = N^2 at
    both walls, so the linear stratification is a steady solution.
"""
from __future__ import annotations

import math

from .config import DIRICHLET, FP64, NEUMANN, Config


def front_dns(nx: int, ny: int, nz: int, re: float, m2: float, n2: float, f: float,
              dt: float, amp: float, precision: int = FP64) -> Config:
    return Config(
        nx=nx, ny=ny, nz=nz, n_th=1, flavor="Front", use_les=False,
        re=re, beta=1.0, lx=2.0 * math.pi, ly=2.0, lz=2.0 * math.pi, nu_v_scale=1.0,
        delta_t=dt, ri=[n2], pr=[1.0], ic_type=4, kick=2.0 * amp, physical_noise=True,
        ro=1.0 / f, delta=1.0, grav_x=0.0, grav_y=1.0, grav_z=0.0, f_type=3,
        u_bc_ymin=DIRICHLET, u_bc_ymin_c1=0.0, w_bc_ymin=DIRICHLET, w_bc_ymin_c1=0.0,
        v_bc_ymin=DIRICHLET, v_bc_ymin_c1=0.0,
        u_bc_ymax=NEUMANN, u_bc_ymax_c1=0.0, w_bc_ymax=NEUMANN, w_bc_ymax_c1=-m2 / f,
        v_bc_ymax=DIRICHLET, v_bc_ymax_c1=0.0,
        th_bc_ymin=[NEUMANN], th_bc_ymin_c1=[n2], th_bc_ymax=[NEUMANN], th_bc_ymax_c1=[n2],
        grid_type=1, grid_cs=1.75, dthdx=[m2], dthdz=[0.0], precision=precision,
        bc_compensation=False,  # w_bc_ymax_c1 above already holds it
    ).derive()


def dns_64(precision: int = FP64) -> Config:
    """Plain DNS, infinite front, 64x65x64: Re 1000, f 1, M^2 0.5, N^2 1 (Ri 4),
    dt 1e-3, perturbation amplitude 1e-3."""
    return (front_dns(64, 65, 64, 1000.0, 0.5, 1.0, 1.0, 1e-3, 1e-3, precision))


def dns_256(precision: int = FP64) -> Config:
    """Plain DNS, same front, 256x129x256: Re 2000, Ri 1 (N^2 = Ri M^4 / f^2 = 0.25),
    dt 5e-4, perturbation amplitude 1e-3."""
    return (front_dns(256, 129, 256, 2000.0, 0.5, 0.25, 1.0, 5e-4, 1e-3, precision))


def small_dns(nx: int = 16, ny: int = 17, nz: int = 16, precision: int = FP64) -> Config:
    """A small grid of the same physics for fast tests."""
    return (front_dns(nx, ny, nz, 1000.0, 0.5, 1.0, 1.0, 1e-3, 1e-3, precision))


def front_les(nx: int, ny: int, nz: int, re: float, dt: float, amp: float, n2: float = 1.0,
              precision: int = FP64, length: float = 8.0 * math.pi) -> Config:
    """Finite front with the AMD closure (les_model_type 5, C_AMD 0.2887 of
    for/les.f90:25): Lx = Lz = `length` (8 pi), f = Ro_inv / delta = 1 with
    Ro_inv = delta = 0.5, so the tanh front of IC type 6 (for/flow.f90:629-646,
    366-385) has width 0.5 and M^2 = 1 at its core, stratified with Ri(1) = N^2,
    in thermal-wind balance; the background gradient dTHdX = 2 Ro_inv / Lx and the
    top Neumann compensation of w come from input_chan (for/parameters.f90:266-273).
    Bottom no-slip, top free-slip, scalar no-flux walls."""
    return Config(
        nx=nx, ny=ny, nz=nz, n_th=1, flavor="Front", use_les=True, les_model_type=5,
        re=re, beta=1.0, lx=length, ly=2.0, lz=length, nu_v_scale=1.0,
        delta_t=dt, ri=[n2], pr=[1.0], ic_type=6, kick=2.0 * amp, physical_noise=True,
        ro=2.0, delta=0.5, grav_x=0.0, grav_y=1.0, grav_z=0.0, f_type=3,
        u_bc_ymin=DIRICHLET, u_bc_ymin_c1=0.0, w_bc_ymin=DIRICHLET, w_bc_ymin_c1=0.0,
        v_bc_ymin=DIRICHLET, v_bc_ymin_c1=0.0,
        u_bc_ymax=NEUMANN, u_bc_ymax_c1=0.0, w_bc_ymax=NEUMANN, w_bc_ymax_c1=0.0,
        v_bc_ymax=DIRICHLET, v_bc_ymax_c1=0.0,
        th_bc_ymin=[NEUMANN], th_bc_ymin_c1=[0.0], th_bc_ymax=[NEUMANN], th_bc_ymax_c1=[0.0],
        grid_type=1, grid_cs=1.75, precision=precision, bc_compensation=True,
    ).derive()


def les_512(precision: int = FP64) -> Config:
    """LES with the AMD closure, finite front, 512x257x512: Re 1e5, N^2 1, dt 2e-4,
    perturbation amplitude 1e-2."""
    return front_les(512, 257, 512, 1e5, 2e-4, 1e-2, precision=precision)


def les_128(precision: int = FP64) -> Config:
    """The same front at the 128x129x128 LES grid, Re 5000."""
    return front_les(128, 129, 128, 5000.0, 2e-4, 1e-2, precision=precision)


def les_1024(precision: int = FP64) -> Config:
    """The front at 1024x385x1024: Re 2e5, N^2 1, dt 1e-4, amplitude 1e-2."""
    return front_les(1024, 385, 1024, 2e5, 1e-4, 1e-2, precision=precision)


NAMED = {"dns_64": dns_64, "dns_256": dns_256, "les_128": les_128, "les_512": les_512, "les_1024": les_1024}

"""ctypes binding of libdgpu_io.so: flow checkpoints (WriteHDF5 / ReadHDF5,
for/phdf5.f90:918-1456) and statistics series (WriteStatH5_Y / WriteHDF5_real,
for/phdf5.f90:7-278) in the reference's HDF5 layout (include/dgpu_io.h).

Physical arrays are numpy views of the reference's Fortran layout
(0:Nx+1, 0:Nzp+1, 0:Nyp+1): C-order shape (Ny+2, Nz+2, Nx+2), float64.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import layout
from . import solver as _solver

IO_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libdgpu_io.so")

_io = None


class FlowTimes(ctypes.Structure):
    """The Timestep attributes of a flow file."""
    _fields_ = [("time", ctypes.c_double), ("save_flow_time", ctypes.c_double),
                ("save_stats_time", ctypes.c_double), ("save_movie_time", ctypes.c_double),
                ("time_step", ctypes.c_int32)]

    def as_dict(self) -> Dict[str, float]:
        return {k: getattr(self, k) for k, _ in self._fields_}


def load():
    """Load libdgpu_io.so (after libdiablo_gpu.so, which it links against)."""
    global _io
    if _io is not None:
        return _io
    _solver.load()
    if not os.path.exists(IO_PATH):
        raise _solver.LibraryMissing(f"{IO_PATH} is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(IO_PATH)
    P = ctypes.c_void_p
    D = ctypes.POINTER(ctypes.c_double)
    I32 = ctypes.c_int32
    T = ctypes.POINTER(FlowTimes)
    s = ctypes.c_char_p
    L.dgpu_h5_write_flow.argtypes = [s, I32, I32, I32, I32, D, D, D, D, D, T, s]
    L.dgpu_h5_read_flow.argtypes = [s, I32, I32, I32, I32, ctypes.c_uint32, D, D, D, D, D,
                                    ctypes.POINTER(I32), T]
    L.dgpu_g2gf.argtypes = [D, I32, I32, I32]
    L.dgpu_g2gf.restype = None
    L.dgpu_gf2g.argtypes = [D, I32, I32, I32]
    L.dgpu_gf2g.restype = None
    L.dgpu_h5_append_profile.argtypes = [s, s, D, I32, ctypes.c_double]
    L.dgpu_h5_append_real.argtypes = [s, s, ctypes.c_double]
    L.dgpu_h5_write_grid.argtypes = [s, D, I32]
    L.dgpu_h5_read_grid.argtypes = [s, D, ctypes.POINTER(I32)]
    L.dgpu_h5_append_plane.argtypes = [s, s, D, I32, I32, ctypes.c_double]
    L.dgpu_write_flow.argtypes = [P, s, T, I32]
    L.dgpu_read_flow.argtypes = [P, s, ctypes.c_uint32, T, ctypes.POINTER(I32)]
    L.dgpu_io_last_error.restype = ctypes.c_char_p
    _io = L
    return L


def _check(rc):
    if rc:
        raise RuntimeError(f"libdgpu_io error {rc}: {load().dgpu_io_last_error().decode()}")


def _dp(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _times(t) -> FlowTimes:
    if isinstance(t, FlowTimes):
        return t
    return FlowTimes(**{k: t.get(k, 0) for k, _ in FlowTimes._fields_})


def write_grid(fname: str, gy: np.ndarray) -> None:
    """grid.h5:/grids/y (create_grid.m's file) from the Ny + 1 values G(2..Ny+2)."""
    g = np.ascontiguousarray(gy, np.float64)
    _check(load().dgpu_h5_write_grid(fname.encode(), _dp(g), len(g)))


def read_grid(fname: str) -> np.ndarray:
    """The values of grid.h5:/grids/y."""
    n = ctypes.c_int32()
    _check(load().dgpu_h5_read_grid(fname.encode(), None, ctypes.byref(n)))
    g = np.zeros(n.value, np.float64)
    _check(load().dgpu_h5_read_grid(fname.encode(), _dp(g), ctypes.byref(n)))
    return g


def physical_shape(nx: int, ny: int, nz: int):
    return (ny + 2, nz + 2, nx + 2)


def g2gf(var: np.ndarray, nx: int, ny: int, nz: int) -> None:
    """u2 from the GY points to the fractional grid, in place (for/tools.f90:299-315)."""
    load().dgpu_g2gf(_dp(var), nx, ny, nz)


def gf2g(var: np.ndarray, nx: int, ny: int, nz: int) -> None:
    """The inverse map with the zero-mass-flux wall rows, in place (for/tools.f90:255-295)."""
    load().dgpu_gf2g(_dp(var), nx, ny, nz)


def write_physical(fname: str, nx: int, ny: int, nz: int, u1, u2f, u3, th: Sequence[np.ndarray] = (),
                   p: Optional[np.ndarray] = None, times=None, date: Optional[str] = None) -> None:
    """WriteHDF5's file from physical fields (u2f: u2 on the fractional grid)."""
    nth = len(th)
    tha = np.ascontiguousarray(np.stack(th)) if nth else None
    t = _times(times or {})
    _check(load().dgpu_h5_write_flow(fname.encode(), nx, ny, nz, nth, _dp(u1), _dp(u2f), _dp(u3), _dp(tha),
                                     _dp(p), ctypes.byref(t), date.encode() if date else None))


def read_physical(fname: str, nx: int, ny: int, nz: int, nth: int, th_mask: Optional[int] = None) -> dict:
    """ReadHDF5's file reading: fields through swapyz, the Timestep times, P if present."""
    shp = physical_shape(nx, ny, nz)
    u1, u2f, u3, p = (np.zeros(shp) for _ in range(4))
    th = np.zeros((max(nth, 1),) + shp)
    has_p = ctypes.c_int32(0)
    t = FlowTimes()
    mask = (1 << nth) - 1 if th_mask is None else th_mask
    _check(load().dgpu_h5_read_flow(fname.encode(), nx, ny, nz, nth, mask, _dp(u1), _dp(u2f), _dp(u3),
                                    _dp(th), _dp(p), ctypes.byref(has_p), ctypes.byref(t)))
    return {"u1": u1, "u2f": u2f, "u3": u3, "th": [th[n] for n in range(nth)],
            "p": p if has_p.value else None, "times": t.as_dict()}


def append_profile(fname: str, gname: str, diag: np.ndarray, time: float) -> None:
    """WriteStatH5_Y: one more sample of a y profile (Nyp values) in group gname."""
    d = np.ascontiguousarray(diag, dtype=np.float64)
    _check(load().dgpu_h5_append_profile(fname.encode(), gname.encode(), _dp(d), d.size, time))


def append_plane(fname: str, gname: str, plane: np.ndarray, time: float) -> None:
    """WriteHDF5_XYplane / XZplane / ZYplane: one more sample of a plane, given in
    the file's C layout [n2][n1] (the Fortran var2d(n1, n2))."""
    d = np.ascontiguousarray(plane, dtype=np.float64)
    n2, n1 = d.shape
    _check(load().dgpu_h5_append_plane(fname.encode(), gname.encode(), _dp(d), n1, n2, time))


def append_real(fname: str, gname: str, value: float) -> None:
    """WriteHDF5_real: one more sample of a scalar in group gname."""
    _check(load().dgpu_h5_append_real(fname.encode(), gname.encode(), float(value)))


def write_flow(solver: "_solver.Solver", fname: str, times, save_pressure: bool = True) -> None:
    """WriteHDF5 of a Solver's device state (which then holds the reference's
    post-write state: u2 through g2gf/gf2g, every field through the transform
    round trip)."""
    t = _times(times)
    _check(load().dgpu_write_flow(solver.h, fname.encode(), ctypes.byref(t), int(bool(save_pressure))))


def read_flow(solver: "_solver.Solver", fname: str, th_mask: Optional[int] = None):
    """ReadHDF5 into a Solver's device state; returns (times dict, read_pressure)."""
    nth = solver.cfg.n_th
    mask = (1 << nth) - 1 if th_mask is None else th_mask
    t = FlowTimes()
    rp = ctypes.c_int32(0)
    _check(load().dgpu_read_flow(solver.h, fname.encode(), mask, ctypes.byref(t), ctypes.byref(rp)))
    return t.as_dict(), bool(rp.value)

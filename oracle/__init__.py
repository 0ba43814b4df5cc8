"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker, and by bench.py's cpu_baseline leg.  The product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        P = ctypes.c_void_p
        D = ctypes.POINTER(ctypes.c_double)
        L.oracle_create.argtypes = [P, ctypes.POINTER(P)]
        L.oracle_destroy.argtypes = [P]
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_set_variant.argtypes = [ctypes.c_int]
        L.oracle_last_error.restype = ctypes.c_char_p
        for name in ("oracle_upload_state", "oracle_download_state"):
            getattr(L, name).argtypes = [P, D, D, D, D, D]
        for name in ("oracle_upload_rk", "oracle_download_rk"):
            getattr(L, name).argtypes = [P, D, D, D, D]
        L.oracle_download_les.argtypes = [P, D, D]
        L.oracle_rk_substep.argtypes = [P, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int, ctypes.c_int]
        L.oracle_rem_div.argtypes = [P]
        L.oracle_filter_chan.argtypes = [P, ctypes.c_int]
        L.oracle_bc_post.argtypes = [P, ctypes.c_double]
        L.oracle_set_variable_dt.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.oracle_courant.argtypes = [P, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, D, ctypes.POINTER(ctypes.c_int)]
        L.oracle_courant_schedule.argtypes = [ctypes.c_double] * 5
        L.oracle_courant_schedule.restype = ctypes.c_double
        L.oracle_courant_dt_state.argtypes = [P, ctypes.c_double]
        L.oracle_courant_dt_state.restype = ctypes.c_double
        L.oracle_fft_xz_to_fourier.argtypes = [P, D, D]
        L.oracle_fft_xz_to_physical.argtypes = [P, D, D]
        L.oracle_wavenumbers.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_double, D, D]
        L.oracle_grid_metrics.argtypes = [ctypes.c_int, D, D, D, D, D]
        L.oracle_cn_matrix.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, D, D, D, D]
        L.oracle_cn_matrix_nut.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, D, D, D, D, D]
        L.oracle_debug_rhs.argtypes = [P, ctypes.c_int, D]
        L.oracle_les_capture.argtypes = [P, D]
        L.oracle_les_override.argtypes = [P, D, D]
        L.oracle_thomas_real.argtypes = [ctypes.c_int, D, D, D, D]
        L.oracle_thomas_complex.argtypes = [ctypes.c_int, D, D, D, D]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """One CPU oracle instance over a Problem's C config, Fortran-layout I/O."""

    def __init__(self, problem, threads: int = 0):
        from diablo3_amd import layout
        self.p = problem
        self.cfg = problem.cfg
        self._layout = layout
        L = lib()
        if threads:
            L.oracle_set_threads(threads)
        h = ctypes.c_void_p()
        rc = L.oracle_create(ctypes.cast(problem.cconfig.ptr, ctypes.c_void_p), ctypes.byref(h))
        if rc:
            raise RuntimeError(L.oracle_last_error().decode())
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def _check(self, rc):
        if rc:
            raise RuntimeError(lib().oracle_last_error().decode())

    def _fo(self, compact):
        return self._layout.to_fortran_fourier(self.cfg, compact)

    def upload_state(self, s):
        from diablo3_amd.ic import FourierState  # noqa: F401
        f = [self._fo(a) for a in (s.cu1, s.cu2, s.cu3, s.cp)]
        th = np.stack([self._fo(t) for t in s.cth]) if s.cth else np.zeros((1,) + f[0].shape, np.complex128)
        self._keep = (f, th)
        self._check(lib().oracle_upload_state(self.h, *[_dp(a.view(np.float64)) for a in f],
                                              _dp(th.view(np.float64))))

    def download_state(self):
        from diablo3_amd.ic import FourierState
        shp = self._layout.fortran_fourier_shape(self.cfg)
        f = [np.zeros(shp, np.complex128) for _ in range(4)]
        th = np.zeros((max(self.cfg.n_th, 1),) + shp, np.complex128)
        self._check(lib().oracle_download_state(self.h, *[_dp(a.view(np.float64)) for a in f],
                                                _dp(th.view(np.float64))))
        c = [self._layout.from_fortran_fourier(self.cfg, a) for a in f]
        cth = [self._layout.from_fortran_fourier(self.cfg, th[n]) for n in range(self.cfg.n_th)]
        return FourierState(c[0], c[1], c[2], c[3], cth)

    def download_rk(self):
        shp = self._layout.fortran_fourier_shape(self.cfg)
        f = [np.zeros(shp, np.complex128) for _ in range(3)]
        th = np.zeros((max(self.cfg.n_th, 1),) + shp, np.complex128)
        self._check(lib().oracle_download_rk(self.h, *[_dp(a.view(np.float64)) for a in f],
                                             _dp(th.view(np.float64))))
        return ([self._layout.from_fortran_fourier(self.cfg, a) for a in f],
                [self._layout.from_fortran_fourier(self.cfg, th[n]) for n in range(self.cfg.n_th)])

    def download_les(self):
        """nu_t and kappa_t[n] as compact physical [j][z][x] arrays."""
        shp = self._layout.fortran_physical_shape(self.cfg)
        nt = np.zeros(shp, np.float64)
        kt = np.zeros((max(self.cfg.n_th, 1),) + shp, np.float64)
        self._check(lib().oracle_download_les(self.h, _dp(nt), _dp(kt)))
        return (self._layout.from_fortran_physical(self.cfg, nt),
                [self._layout.from_fortran_physical(self.cfg, kt[n]) for n in range(self.cfg.n_th)])

    def rk_substep(self, rk_step, delta_t=None, time=0.0, time_step=1, create_new_flow=1):
        dt = self.cfg.delta_t if delta_t is None else delta_t
        self._check(lib().oracle_rk_substep(self.h, rk_step, dt, time, time_step, int(create_new_flow)))

    def step(self, delta_t=None, time=0.0, time_step=1, create_new_flow=1):
        for rk in (1, 2, 3):
            self.rk_substep(rk, delta_t, time, time_step, create_new_flow)

    def rem_div(self):
        self._check(lib().oracle_rem_div(self.h))

    def filter_chan(self, n=0):
        self._check(lib().oracle_filter_chan(self.h, n))

    def set_variable_dt(self, variable_dt, update_dt=1, cfl=0.5):
        self._check(lib().oracle_set_variable_dt(self.h, int(variable_dt), update_dt, cfl))

    def courant(self, time, time_limit, save_stats_time, save_flow_time):
        """(delta_t, ran) of the courant call inside the last rk_step 3."""
        dt, ran = ctypes.c_double(), ctypes.c_int()
        self._check(lib().oracle_courant(self.h, time, time_limit, save_stats_time, save_flow_time,
                                         ctypes.byref(dt), ctypes.byref(ran)))
        return dt.value, bool(ran.value)

    def courant_dt_state(self, cfl):
        return lib().oracle_courant_dt_state(self.h, cfl)

    def bc_post(self, time=0.0):
        self._check(lib().oracle_bc_post(self.h, time))

    def fft_to_fourier(self, phys_compact):
        pf = self._layout.to_fortran_physical(self.cfg, phys_compact)
        out = np.zeros(self._layout.fortran_fourier_shape(self.cfg), np.complex128)
        self._check(lib().oracle_fft_xz_to_fourier(self.h, _dp(pf), _dp(out.view(np.float64))))
        return self._layout.from_fortran_fourier(self.cfg, out)

    def fft_to_physical(self, four_compact):
        ff = self._fo(four_compact)
        out = np.zeros(self._layout.fortran_physical_shape(self.cfg), np.float64)
        self._check(lib().oracle_fft_xz_to_physical(self.h, _dp(ff.view(np.float64)), _dp(out)))
        return self._layout.from_fortran_physical(self.cfg, out)

    def cn_matrix(self, which, rk_step, delta_t=None):
        n2 = self.cfg.ny2
        dt = self.cfg.delta_t if delta_t is None else delta_t
        a, b, c, v = (np.zeros(n2) for _ in range(4))
        self._check(lib().oracle_cn_matrix(self.h, which, rk_step, dt, _dp(a), _dp(b), _dp(c), _dp(v)))
        return a, b, c, v

    def cn_matrix_nut(self, which, rk_step, nut, delta_t=None):
        """CN system of one column with the eddy-viscosity column nut (rows 0..Nyp+1)."""
        n2 = self.cfg.ny2
        dt = self.cfg.delta_t if delta_t is None else delta_t
        nut = np.ascontiguousarray(nut, dtype=np.float64)
        assert nut.shape == (n2,)
        a, b, c, v = (np.zeros(n2) for _ in range(4))
        self._check(lib().oracle_cn_matrix_nut(self.h, which, rk_step, dt, _dp(nut), _dp(a), _dp(b),
                                               _dp(c), _dp(v)))
        return a, b, c, v

    def les_capture(self, on=True):
        """Capture array [ny2][nz][nx][16] the next LES substeps fill (see diablo_oracle.h)."""
        if not on:
            self._check(lib().oracle_les_capture(self.h, None))
            self._cap = None
            return None
        self._cap = np.zeros((self.cfg.ny2, self.cfg.nz, self.cfg.nx, 16), np.float64)
        self._check(lib().oracle_les_capture(self.h, _dp(self._cap)))
        return self._cap

    def les_override(self, cgy, kgy):
        """GY coefficients [ny2][nz][nx] used as les_model_type 6 uses them (None: off)."""
        if cgy is None:
            self._ov = None
            self._check(lib().oracle_les_override(self.h, None, None))
            return
        self._ov = (np.ascontiguousarray(cgy, np.float64), np.ascontiguousarray(kgy, np.float64))
        self._check(lib().oracle_les_override(self.h, _dp(self._ov[0]), _dp(self._ov[1])))

    def debug_rhs(self, which):
        """The physical CN right-hand side of field `which` (0 u1, 1 u2, 2 u3, 3+n scalar n)
        left by the last substep, as a compact [j][z][x] array."""
        out = np.zeros(self._layout.fortran_physical_shape(self.cfg), np.float64)
        self._check(lib().oracle_debug_rhs(self.h, which, _dp(out)))
        return self._layout.from_fortran_physical(self.cfg, out)


def wavenumbers(nx, nz, lx, lz):
    kx = np.zeros(nx // 2)
    kz = np.zeros(2 * (nz // 3) + 1)
    lib().oracle_wavenumbers(nx, nz, lx, lz, _dp(kx), _dp(kz))
    return kx, kz


def grid_metrics(gy_file):
    ny = len(gy_file) - 1
    g = np.ascontiguousarray(gy_file, dtype=np.float64)
    out = [np.zeros(ny + 2) for _ in range(4)]
    lib().oracle_grid_metrics(ny, _dp(g), *[_dp(a) for a in out])
    return dict(zip(("gy", "gyf", "dy", "dyf"), out))


def thomas_real(a, b, c, g):
    a, b, c, g = (np.array(x, dtype=np.float64) for x in (a, b, c, g))
    lib().oracle_thomas_real(len(a) - 2, _dp(a), _dp(b), _dp(c), _dp(g))
    return g


def thomas_complex(a, b, c, g):
    a, b, c = (np.array(x, dtype=np.float64) for x in (a, b, c))
    g = np.array(g, dtype=np.complex128)
    lib().oracle_thomas_complex(len(a) - 2, _dp(a), _dp(b), _dp(c), _dp(g.view(np.float64)))
    return g

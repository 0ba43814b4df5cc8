"""ctypes binding of libdiablo_gpu.so — the drop-in for ``call rk_chan_1``.

The binding is thin on purpose: every array crosses the C ABI in the reference's
Fortran layout (see include/diablo_gpu.h), exactly as the Fortran driver would
pass it through iso_c_binding.  There is no fallback: if the library is missing or
no GPU is present, construction raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Optional

import numpy as np

from . import layout
from .ic import FourierState

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdiablo_gpu.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "diablo_gpu.h")

_lib = None


class LibraryMissing(RuntimeError):
    pass


def load():
    """Load libdiablo_gpu.so (raises LibraryMissing if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    D = ctypes.POINTER(ctypes.c_double)
    L.dgpu_create.argtypes = [P, ctypes.POINTER(P)]
    L.dgpu_destroy.argtypes = [P]
    for name in ("dgpu_upload_state", "dgpu_download_state"):
        getattr(L, name).argtypes = [P, D, D, D, D, D]
    for name in ("dgpu_upload_rk", "dgpu_download_rk"):
        getattr(L, name).argtypes = [P, D, D, D, D]
    L.dgpu_rk_substep.argtypes = [P, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_int32, ctypes.c_int32]
    L.dgpu_step.argtypes = [P, ctypes.c_double, ctypes.c_double, ctypes.c_int32, ctypes.c_int32]
    L.dgpu_rem_div.argtypes = [P]
    L.dgpu_set_graph.argtypes = [P, ctypes.c_int32]
    L.dgpu_timing_every.argtypes = [P, ctypes.c_int32]
    L.dgpu_poisson_p.argtypes = [P]
    L.dgpu_filter_chan.argtypes = [P, ctypes.c_int32]
    L.dgpu_fft_xz_to_fourier.argtypes = [P, D, D]
    L.dgpu_fft_xz_to_physical.argtypes = [P, D, D]
    L.dgpu_fft_xz_to_fourier_noalloc.argtypes = [P, P, P]
    L.dgpu_synchronize.argtypes = [P]
    L.dgpu_last_step_ms.argtypes = [P, D]
    L.dgpu_last_error.restype = ctypes.c_char_p
    L.dgpu_timing_enable.argtypes = [P, ctypes.c_int32]
    L.dgpu_timing_read.argtypes = [P, ctypes.c_int32, D, ctypes.POINTER(ctypes.c_int64),
                                   ctypes.POINTER(ctypes.c_char_p)]
    L.dgpu_download_les.argtypes = [P, D, D]
    L.dgpu_stats.argtypes = [P, ctypes.c_double, D]
    L.dgpu_set_variable_dt.argtypes = [P, ctypes.c_int32, ctypes.c_int32, ctypes.c_double]
    L.dgpu_courant.argtypes = [P, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                               D, ctypes.POINTER(ctypes.c_int32)]
    L.dgpu_courant_physical.argtypes = [P, D, D, D]
    L.dgpu_stats_doubles.argtypes = [P]
    L.dgpu_stats_doubles.restype = ctypes.c_int64
    L.dgpu_stats_zstar.argtypes = [P, ctypes.c_double, D]
    L.dgpu_stats_zstar_doubles.argtypes = [P]
    L.dgpu_stats_zstar_doubles.restype = ctypes.c_int64
    L.dgpu_stats_movie.argtypes = [P, ctypes.c_double, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, D, D]
    L.dgpu_stats_movie_sizes.argtypes = [P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
    L.dgpu_get_sizes.argtypes = [P] + [ctypes.POINTER(ctypes.c_int32)] * 4
    L.dgpu_field_doubles.argtypes = [P]
    L.dgpu_field_doubles.restype = ctypes.c_int64
    F = ctypes.POINTER(ctypes.c_float)
    I64 = ctypes.c_int64
    L.dgpu_mlp_create.argtypes = [F, ctypes.c_uint64, ctypes.POINTER(P)]
    L.dgpu_mlp_destroy.argtypes = [P]
    L.dgpu_mlp_set_params.argtypes = [P, F]
    L.dgpu_mlp_get_params.argtypes = [P, F]
    L.dgpu_mlp_set_norm.argtypes = [P, F]
    L.dgpu_mlp_set_data.argtypes = [P, F, F, I64]
    L.dgpu_mlp_grad.argtypes = [P, I64, F, F, F]
    L.dgpu_mlp_adam_step.argtypes = [P, F, ctypes.c_float]
    L.dgpu_mlp_train.argtypes = [P, ctypes.c_int32, ctypes.c_int32, ctypes.c_float, ctypes.c_uint64, F, F]
    L.dgpu_mlp_predict.argtypes = [P, F, I64, F]
    L.dgpu_set_closure.argtypes = [P, F, F]
    L.dgpu_box_filter.argtypes = [P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, D, D, D]
    L.dgpu_closure_counts.argtypes = [P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
    L.dgpu_les_features.argtypes = [P, ctypes.c_double, F, F, F, ctypes.POINTER(ctypes.c_int32),
                                    ctypes.POINTER(I64)]
    _lib = L
    return L


# headers whose functions libdiablo_gpu.so exports (include/dgpu_io.h belongs to
# libdgpu_io.so)
LIB_HEADERS = ("diablo_gpu.h", "dgpu_closure.h")


def exported_symbols(headers=LIB_HEADERS) -> List[str]:
    """Names of the functions the given headers under include/ declare."""
    import re
    names = set()
    for h in headers:
        path = os.path.join(os.path.dirname(HEADER_PATH), h)
        names |= set(re.findall(r"\b(dgpu_[a-z0-9_]+)\s*\(", open(path).read()))
    return sorted(names)


def _dp(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Solver:
    """One device-resident channel state; methods mirror the reference routines."""

    def __init__(self, problem):
        L = load()
        self.p = problem
        self.cfg = problem.cfg
        h = ctypes.c_void_p()
        rc = L.dgpu_create(ctypes.cast(problem.cconfig.ptr, ctypes.c_void_p), ctypes.byref(h))
        if rc:
            raise RuntimeError(f"dgpu_create failed ({rc}): {L.dgpu_last_error().decode()}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            load().dgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise RuntimeError(f"libdiablo_gpu error {rc}: {load().dgpu_last_error().decode()}")

    def _fshape(self):
        return layout.fortran_fourier_shape(self.cfg)

    # ---- state ----
    def upload_state(self, s: FourierState):
        f = [layout.to_fortran_fourier(self.cfg, a) for a in (s.cu1, s.cu2, s.cu3, s.cp)]
        th = (np.stack([layout.to_fortran_fourier(self.cfg, t) for t in s.cth]) if s.cth
              else np.zeros((1,) + self._fshape(), np.complex128))
        self._check(load().dgpu_upload_state(self.h, *[_dp(a.view(np.float64)) for a in f],
                                             _dp(th.view(np.float64))))

    def download_state(self) -> FourierState:
        shp = self._fshape()
        f = [np.zeros(shp, np.complex128) for _ in range(4)]
        th = np.zeros((max(self.cfg.n_th, 1),) + shp, np.complex128)
        self._check(load().dgpu_download_state(self.h, *[_dp(a.view(np.float64)) for a in f],
                                               _dp(th.view(np.float64))))
        c = [layout.from_fortran_fourier(self.cfg, a) for a in f]
        return FourierState(c[0], c[1], c[2], c[3],
                            [layout.from_fortran_fourier(self.cfg, th[n]) for n in range(self.cfg.n_th)])

    def download_rk(self):
        shp = self._fshape()
        f = [np.zeros(shp, np.complex128) for _ in range(3)]
        th = np.zeros((max(self.cfg.n_th, 1),) + shp, np.complex128)
        self._check(load().dgpu_download_rk(self.h, *[_dp(a.view(np.float64)) for a in f],
                                            _dp(th.view(np.float64))))
        return ([layout.from_fortran_fourier(self.cfg, a) for a in f],
                [layout.from_fortran_fourier(self.cfg, th[n]) for n in range(self.cfg.n_th)])

    def download_les(self):
        """nu_t and kappa_t[n] of the last LES substep, compact physical [j][z][x]."""
        shp = layout.fortran_physical_shape(self.cfg)
        nt = np.zeros(shp, np.float64)
        kt = np.zeros((max(self.cfg.n_th, 1),) + shp, np.float64)
        self._check(load().dgpu_download_les(self.h, _dp(nt), _dp(kt)))
        return (layout.from_fortran_physical(self.cfg, nt),
                [layout.from_fortran_physical(self.cfg, kt[n]) for n in range(self.cfg.n_th)])

    def set_closure(self, params: np.ndarray, norm: np.ndarray):
        """The learned closure's network for les_model_type 6 (dgpu_set_closure)."""
        p = np.ascontiguousarray(params, dtype=np.float32)
        n = np.ascontiguousarray(norm, dtype=np.float32)
        F = ctypes.POINTER(ctypes.c_float)
        self._check(load().dgpu_set_closure(self.h, p.ctypes.data_as(F), n.ctypes.data_as(F)))

    def les_features(self, time: float = 0.0):
        """dgpu_les_features: (X [n][16], scale [n], amd [n][2], inside [n])."""
        L = load()
        F = ctypes.POINTER(ctypes.c_float)
        n = ctypes.c_int64()
        self._check(L.dgpu_les_features(self.h, time, None, None, None, None, ctypes.byref(n)))
        X = np.zeros((n.value, 16), np.float32)
        sc = np.zeros(n.value, np.float32)
        am = np.zeros((n.value, 2), np.float32)
        ins = np.zeros(n.value, np.int32)
        self._check(L.dgpu_les_features(self.h, time, X.ctypes.data_as(F), sc.ctypes.data_as(F),
                                        am.ctypes.data_as(F), ins.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                        ctypes.byref(n)))
        return X, sc, am, ins

    def closure_counts(self):
        """dgpu_closure_counts: (points whose coefficient came from the network,
        points the closure evaluated) in the last LES substep."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        self._check(load().dgpu_closure_counts(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def box_filter(self, fx: int, fz: int, mf: np.ndarray, mg: np.ndarray) -> np.ndarray:
        """dgpu_box_filter: the state filtered onto a grid coarser by fx, fz in x, z
        and with the LES rows of the y weights mf, mg ([nl+2][Ny+2]): the 14
        fields of sgs_data.BOX_FIELDS as [14][nl+2][Nz/fz][Nx/fx]."""
        cfg = self.cfg
        mf = np.ascontiguousarray(mf, np.float64)
        mg = np.ascontiguousarray(mg, np.float64)
        assert mf.shape == mg.shape and mf.shape[1] == cfg.ny2
        nl2 = mf.shape[0]
        out = np.zeros((14, nl2, cfg.nz // fz, cfg.nx // fx), np.float64)
        self._check(load().dgpu_box_filter(self.h, fx, fz, nl2 - 2, _dp(mf), _dp(mg), _dp(out)))
        return out

    def set_variable_dt(self, variable_dt: bool, update_dt: int = 1, cfl: float = 0.5):
        self._check(load().dgpu_set_variable_dt(self.h, int(variable_dt), update_dt, cfl))

    def courant(self, time, time_limit, save_stats_time, save_flow_time):
        """(delta_t, ran) of the courant evaluation in the last rk_step 3."""
        dt, ran = ctypes.c_double(), ctypes.c_int32()
        self._check(load().dgpu_courant(self.h, time, time_limit, save_stats_time, save_flow_time,
                                        ctypes.byref(dt), ctypes.byref(ran)))
        return dt.value, bool(ran.value)

    def stats(self, time: float = 0.0):
        """save_stats_chan's energy-budget profiles: {name: (Ny+2,) array}, 'u1z_x0'."""
        from .stats import PROFILES, PER_SCALAR
        n = load().dgpu_stats_doubles(ctypes.cast(self.p.cconfig.ptr, ctypes.c_void_p))
        out = np.zeros(n, np.float64)
        self._check(load().dgpu_stats(self.h, time, _dp(out)))
        n2 = self.cfg.ny2
        names = PROFILES + [f"{b}{k + 1:02d}" for k in range(self.cfg.n_th) for b in PER_SCALAR]
        res = {name: out[i * n2:(i + 1) * n2].copy() for i, name in enumerate(names)}
        k = len(names) * n2
        res["u1z_x0"] = float(out[k])
        res["FTx_uu"] = out[k + 1:k + 1 + self.cfg.nx // 2].copy()
        return res

    def stats_zstar(self, time: float = 0.0):
        """compute_BPE and Bin_Ystar_and_Write of save_stats_chan (scalar 1):
        {'th1PDF', 'th1bin', 'BPE', 'shear2_zstar', 'epsilon_zstar', 'TKE_zstar',
        'KE_zstar', 'thw_zstar01', 'chi_zstar01', ...}."""
        n = load().dgpu_stats_zstar_doubles(ctypes.cast(self.p.cconfig.ptr, ctypes.c_void_p))
        if n < 0:
            raise RuntimeError("z* statistics need a scalar (n_th >= 1)")
        out = np.zeros(n, np.float64)
        self._check(load().dgpu_stats_zstar(self.h, time, _dp(out)))
        nb = self.cfg.nx
        res = {"th1PDF": out[:nb].copy(), "th1bin": out[nb:2 * nb + 1].copy(), "BPE": float(out[2 * nb + 1])}
        names = ["shear2_zstar", "epsilon_zstar", "TKE_zstar", "KE_zstar"]
        names += [f"{b}{k + 1:02d}" for k in range(self.cfg.n_th) for b in ("thw_zstar", "chi_zstar")]
        base = 2 * nb + 2
        for i, name in enumerate(names):
            res[name] = out[base + i * nb:base + (i + 1) * nb].copy()
        return res

    def stats_movie(self, time: float, nxm: int, nym: int, nzm: int):
        """The movie output of save_stats_chan (dgpu_stats_movie): (planes, slices),
        each {group name of the reference: array in the file's C layout}.  planes:
        the mean_xz.h5 z-mean planes [Nyp][Nx]; slices: the movie.h5 slices
        <f>_xz [Nyp][Nx] at z = nzm, <f>_xy [Nz][Nx] at row nym (absent when
        nym = -1), <f>_yz [Nyp][Nz] at x = nxm."""
        from .stats_io import XZ_PLANES, movie_fields
        cfg = self.cfg
        npl, nsl = ctypes.c_int64(), ctypes.c_int64()
        self._check(load().dgpu_stats_movie_sizes(ctypes.cast(self.p.cconfig.ptr, ctypes.c_void_p),
                                                  ctypes.byref(npl), ctypes.byref(nsl)))
        pl = np.zeros(npl.value, np.float64)
        sl = np.zeros(nsl.value, np.float64)
        self._check(load().dgpu_stats_movie(self.h, time, nxm, nym, nzm, _dp(pl), _dp(sl)))
        nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
        planes = {name: pl[i * ny * nx:(i + 1) * ny * nx].reshape(ny, nx).copy() for i, name in enumerate(XZ_PLANES)}
        per = ny * nx + nz * nx + ny * nz
        slices = {}
        for i, name in enumerate(movie_fields(cfg.n_th)):
            f = sl[i * per:(i + 1) * per]
            slices[name + "_xz"] = f[:ny * nx].reshape(ny, nx).copy()
            if nym >= 1:
                slices[name + "_xy"] = f[ny * nx:ny * nx + nz * nx].reshape(nz, nx).copy()
            slices[name + "_yz"] = f[ny * nx + nz * nx:].reshape(ny, nz).copy()
        return planes, slices

    # ---- time advance ----
    def rk_substep(self, rk_step: int, delta_t: Optional[float] = None, time: float = 0.0,
                   time_step: int = 1, create_new_flow: bool = True):
        dt = self.cfg.delta_t if delta_t is None else delta_t
        self._check(load().dgpu_rk_substep(self.h, rk_step, dt, time, time_step, int(create_new_flow)))

    def step(self, delta_t: Optional[float] = None, time: float = 0.0, time_step: int = 1,
             create_new_flow: bool = True):
        dt = self.cfg.delta_t if delta_t is None else delta_t
        self._check(load().dgpu_step(self.h, dt, time, time_step, int(create_new_flow)))

    def filter_chan(self, n: int = 0):
        """filter_chan(n+1) of the reference on the device state (dgpu_filter_chan)."""
        self._check(load().dgpu_filter_chan(self.h, n))

    def set_graph(self, enable: bool):
        """dgpu_step as a captured HIP graph (default) or as direct launches."""
        self._check(load().dgpu_set_graph(self.h, int(bool(enable))))

    def rem_div(self):
        self._check(load().dgpu_rem_div(self.h))

    def poisson_p(self):
        """poisson_p_chan of the reference (for/solver.f90:1382-1560): the pressure
        recomputed from the device velocity and scalars (dgpu_poisson_p)."""
        self._check(load().dgpu_poisson_p(self.h))

    def synchronize(self):
        self._check(load().dgpu_synchronize(self.h))

    def last_step_ms(self) -> float:
        v = ctypes.c_double()
        self._check(load().dgpu_last_step_ms(self.h, ctypes.byref(v)))
        return v.value

    # ---- measurement ----
    def timing(self, enable: bool):
        self._check(load().dgpu_timing_enable(self.h, int(enable)))

    def timing_every(self, every: int):
        """Record the kernel events on every `every`-th step only (dgpu_timing_every)."""
        self._check(load().dgpu_timing_every(self.h, int(every)))

    def timing_read(self):
        """{kernel name: (total device ms, launches)} since timing(True), launched kernels only."""
        out = {}
        for k in range(12):
            ms, n, name = ctypes.c_double(), ctypes.c_int64(), ctypes.c_char_p()
            self._check(load().dgpu_timing_read(self.h, k, ctypes.byref(ms), ctypes.byref(n),
                                                ctypes.byref(name)))
            if n.value:
                out[name.value.decode()] = (ms.value, n.value)
        return out

    # ---- transforms ----
    def fft_to_fourier(self, phys_compact: np.ndarray) -> np.ndarray:
        pf = layout.to_fortran_physical(self.cfg, phys_compact)
        out = np.zeros(self._fshape(), np.complex128)
        self._check(load().dgpu_fft_xz_to_fourier(self.h, _dp(pf), _dp(out.view(np.float64))))
        return layout.from_fortran_fourier(self.cfg, out)

    def fft_to_physical(self, four_compact: np.ndarray) -> np.ndarray:
        ff = layout.to_fortran_fourier(self.cfg, four_compact)
        out = np.zeros(layout.fortran_physical_shape(self.cfg), np.float64)
        self._check(load().dgpu_fft_xz_to_physical(self.h, _dp(ff.view(np.float64)), _dp(out)))
        return layout.from_fortran_physical(self.cfg, out)

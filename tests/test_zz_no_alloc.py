"""dgpu_fft_xz_to_fourier_noalloc through the C ABI: the allocation-free form of
dgpu_fft_xz_to_fourier on device arrays, its intermediates in the library's
static device scratch.

Against the existing entry point on the same input (bit-exact: the same
transform kernels), on two calls in a row on one GPU, the second on a smaller
grid, so that the second call reuses scratch the first one filled; in place
(phys and four one device buffer, as the reference's u/cu views); rows past
2 Nkz are left untouched; free device memory is the same before and after a
call; a grid whose intermediates exceed the scratch is rejected with status 2.
The test's own device buffers come from the HIP runtime the library links
(one runtime in the process).
"""
import ctypes

import numpy as np
import pytest

from diablo3_amd import configs, layout
from diablo3_amd.problem import Problem
from diablo3_amd.solver import Solver, load

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    """The libamdhip64 that libdiablo_gpu.so is linked against, as mapped in this process."""
    load()
    with open("/proc/self/maps") as f:
        paths = {ln.split()[-1] for ln in f if "libamdhip64.so" in ln and ln.split()[-1].startswith("/")}
    lib = next((p for p in sorted(paths) if p.startswith("/opt/rocm")), None) or sorted(paths)[0]
    hip = ctypes.CDLL(lib)
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


class DevBuf:
    def __init__(self, hip, a: np.ndarray):
        self.hip, self.n = hip, a.nbytes
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), self.n) == 0
        assert hip.hipMemcpy(self.p, a.ctypes.data, self.n, H2D) == 0

    def get(self, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, self.n, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _free_bytes(hip):
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _host_path(g, pf, init):
    out = init.copy()
    rc = load().dgpu_fft_xz_to_fourier(g.h, pf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                       out.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0, load().dgpu_last_error()
    return out


def _noalloc(g, phys, four):
    return load().dgpu_fft_xz_to_fourier_noalloc(g.h, phys.p, four.p)


SENT = 7.0 - 3.0j


def _case(hip, nx, ny, nz, seed):
    cfg = configs.small_dns(nx, ny, nz)
    g = Solver(Problem(cfg))
    rng = np.random.default_rng(seed)
    phys = rng.standard_normal((cfg.ny2, cfg.nz, cfg.nx))
    pf = np.ascontiguousarray(layout.to_fortran_physical(cfg, phys))
    sentinel = np.full(layout.fortran_fourier_shape(cfg), SENT)
    want = _host_path(g, pf, sentinel)
    dp, df = DevBuf(hip, pf), DevBuf(hip, sentinel)
    hip.hipDeviceSynchronize()
    free0 = _free_bytes(hip)
    rc = _noalloc(g, dp, df)
    assert rc == 0, load().dgpu_last_error()
    free1 = _free_bytes(hip)
    got = df.get(sentinel)
    dp.free()
    df.free()
    return g, pf, want, got, free0, free1


def test_noalloc_matches_host_entry_point():
    hip = _hip()
    print("noalloc: 32x33x32 then 16x17x16 on device 0", flush=True)
    ga, pfa, want, got, _, _ = _case(hip, 32, 33, 32, seed=1)
    assert np.array_equal(got, want)
    assert np.all(got[:, ga.cfg.kzr:, :] == SENT)
    # second call, smaller grid, on the scratch the first call filled
    gb, pfb, want_b, got_b, free0, free1 = _case(hip, 16, 17, 16, seed=2)
    assert np.array_equal(got_b, want_b)
    assert np.all(got_b[:, gb.cfg.kzr:, :] == SENT)
    assert free1 == free0, (free0, free1)
    # in place: one device buffer as phys and four
    buf = DevBuf(hip, pfb)
    assert _noalloc(gb, buf, buf) == 0, load().dgpu_last_error()
    shp = layout.fortran_fourier_shape(gb.cfg)
    want_alias = _host_path(gb, pfb, pfb.view(np.complex128).reshape(shp))
    assert np.array_equal(buf.get(pfb).view(np.complex128).reshape(shp), want_alias)
    buf.free()
    # the first handle again, after the smaller call
    dp, df = DevBuf(hip, pfa), DevBuf(hip, np.full(want.shape, SENT))
    assert _noalloc(ga, dp, df) == 0, load().dgpu_last_error()
    assert np.array_equal(df.get(want), want)
    dp.free()
    df.free()
    ga.close()
    gb.close()


def test_noalloc_rejects_grids_beyond_the_scratch():
    hip = _hip()
    cfg = configs.small_dns(256, 65, 256)
    g = Solver(Problem(cfg))
    buf = DevBuf(hip, np.zeros(layout.fortran_physical_shape(cfg)))
    rc = _noalloc(g, buf, buf)
    msg = load().dgpu_last_error().decode()
    print("256x65x256:", rc, msg, flush=True)
    assert rc == 2 and "scratch" in msg
    buf.free()
    g.close()

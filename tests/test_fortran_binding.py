"""The Fortran (iso_c_binding) side of the boundary: diablo3_amd/fortran/.

CPU: the bind(C) derived type has the C struct's size and member offsets.
GPU: a Fortran host (dgpu_drive) stepping the library gives exactly what the
ctypes host gives for the same input, and matches the oracle.
"""
import ctypes
import os
import shutil
import struct
import subprocess
import tempfile

import numpy as np
import pytest

from diablo3_amd import configs, layout
from diablo3_amd.config import DgpuConfig
from diablo3_amd.problem import Problem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "diablo3_amd", "fortran")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
MAGIC = 1146114389

needs_flang = pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not installed")


@needs_flang
def test_fortran_type_matches_c_struct():
    prog = """
program t
  use, intrinsic :: iso_c_binding
  use diablo_gpu
  type(dgpu_config), target :: c
  print '(6I8)', c_sizeof(c), &
    transfer(c_loc(c%lx), 0_c_intptr_t) - transfer(c_loc(c), 0_c_intptr_t), &
    transfer(c_loc(c%th_bc_ymin_c1), 0_c_intptr_t) - transfer(c_loc(c), 0_c_intptr_t), &
    transfer(c_loc(c%pr), 0_c_intptr_t) - transfer(c_loc(c), 0_c_intptr_t), &
    transfer(c_loc(c%gy), 0_c_intptr_t) - transfer(c_loc(c), 0_c_intptr_t), &
    transfer(c_loc(c%kz), 0_c_intptr_t) - transfer(c_loc(c), 0_c_intptr_t)
end program
"""
    with tempfile.TemporaryDirectory() as d:
        shutil.copy(os.path.join(FDIR, "diablo_gpu.f90"), d)
        open(os.path.join(d, "t.f90"), "w").write(prog)
        subprocess.run([FLANG, "-c", "diablo_gpu.f90"], cwd=d, check=True)
        libdir = os.path.join(ROOT, "diablo3_amd")
        if not os.path.exists(os.path.join(libdir, "libdiablo_gpu.so")):
            subprocess.run(["make", "-s", "-C", os.path.join(libdir, "csrc")], check=True)
        subprocess.run([FLANG, "t.f90", "diablo_gpu.o", "-L" + libdir, "-ldiablo_gpu",
                        "-Wl,-rpath," + libdir, "-o", "t"], cwd=d, check=True)
        out = subprocess.run([os.path.join(d, "t")], check=True, capture_output=True, text=True).stdout
    got = [int(x) for x in out.split()]
    exp = [ctypes.sizeof(DgpuConfig), DgpuConfig.lx.offset, DgpuConfig.th_bc_ymin_c1.offset,
           DgpuConfig.pr.offset, DgpuConfig.gy.offset, DgpuConfig.kz.offset]
    assert got == exp


def write_drive_input(path, p, state, nsteps, create_new_flow=1):
    cfg = p.cfg
    with open(path, "wb") as f:
        f.write(struct.pack("<iidi", MAGIC, nsteps, cfg.delta_t, create_new_flow))
        f.write(bytes(p.cconfig.struct))
        for k in ("gy", "gyf", "dy", "dyf"):
            f.write(np.ascontiguousarray(p.met[k], np.float64).tobytes())
        f.write(np.ascontiguousarray(p.kx, np.float64).tobytes())
        f.write(np.ascontiguousarray(p.kz, np.float64).tobytes())
        for a in (state.cu1, state.cu2, state.cu3, state.cp, *state.cth):
            f.write(layout.to_fortran_fourier(cfg, a).tobytes())


def read_drive_output(path, p):
    from diablo3_amd.ic import FourierState
    cfg = p.cfg
    shp = layout.fortran_fourier_shape(cfg)
    n = int(np.prod(shp))
    raw = np.fromfile(path, dtype=np.complex128)
    arrs = [layout.from_fortran_fourier(cfg, raw[m * n:(m + 1) * n].reshape(shp)) for m in range(4 + cfg.n_th)]
    return FourierState(arrs[0], arrs[1], arrs[2], arrs[3], arrs[4:])


@pytest.mark.gpu
def test_fortran_host_matches_ctypes_host_and_oracle():
    import oracle
    from diablo3_amd.solver import Solver
    exe = os.path.join(FDIR, "dgpu_drive")
    if not os.path.exists(exe):
        pytest.skip("dgpu_drive not built (no flang at build time)")
    cfg = configs.small_dns(32, 33, 32)
    p = Problem(cfg)
    o = oracle.Oracle(p)
    o.upload_state(p.initial_state(seed=9, theta_noise=1e-3))
    o.rem_div()
    s0 = o.download_state()
    nsteps = 3
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        write_drive_input(fin, p, s0, nsteps)
        r = subprocess.run([exe, fin, fout], capture_output=True, text=True, timeout=120)
        print(r.stdout, r.stderr, flush=True)
        assert r.returncode == 0
        sf = read_drive_output(fout, p)
    g = Solver(p)
    g.upload_state(s0)
    t = 0.0
    for n in range(nsteps):
        for rk in (1, 2, 3):
            g.rk_substep(rk, time=t, time_step=n + 1)
            o.rk_substep(rk, time=t, time_step=n + 1)
        t += cfg.delta_t
    sg, so = g.download_state(), o.download_state()
    for a, b, c in ((sf.cu1, sg.cu1, so.cu1), (sf.cu2, sg.cu2, so.cu2), (sf.cu3, sg.cu3, so.cu3),
                    (sf.cth[0], sg.cth[0], so.cth[0])):
        assert np.array_equal(a, b)
        assert np.linalg.norm(a - c) <= 1e-11 * np.linalg.norm(c)
    assert np.array_equal(sf.cp, sg.cp)

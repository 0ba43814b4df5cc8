"""The C ABI library loads without a GPU and exports every function the header declares."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from diablo3_amd import configs, solver
from diablo3_amd.config import DgpuConfig
from diablo3_amd.problem import Problem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(solver.LIB_PATH):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "diablo3_amd", "csrc")], check=True)
    return ctypes.CDLL(solver.LIB_PATH)


def test_every_declared_symbol_is_exported(lib):
    names = solver.exported_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n


def test_io_library_exports_its_header(lib):
    """libdgpu_io.so (checkpoints, statistics files) loads without a GPU and exports
    every function include/dgpu_io.h declares; every header is covered by one of
    the two libraries."""
    import glob
    from diablo3_amd import flow_io
    if not os.path.exists(flow_io.IO_PATH):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "diablo3_amd", "csrc")], check=True)
    io = flow_io.load()
    names = solver.exported_symbols(("dgpu_io.h",))
    assert len(names) >= 8
    for n in names:
        assert hasattr(io, n), n
    headers = {os.path.basename(p) for p in glob.glob(os.path.join(ROOT, "include", "*.h"))}
    assert headers == set(solver.LIB_HEADERS) | {"dgpu_io.h"}


def test_config_struct_layout_matches_header():
    # offsets the C compiler gives dgpu_config, from a tiny host program
    src = os.path.join(ROOT, "include", "diablo_gpu.h")
    prog = r'''
#include <stddef.h>
#include <stdio.h>
#include "%s"
int main(void){
 printf("%%zu %%zu %%zu %%zu %%zu %%zu\n", sizeof(dgpu_config), offsetof(dgpu_config, lx),
        offsetof(dgpu_config, th_bc_ymin_c1), offsetof(dgpu_config, pr), offsetof(dgpu_config, gy),
        offsetof(dgpu_config, kz));
 return 0;}
''' % src
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(prog)
        subprocess.run(["gcc", "-o", os.path.join(d, "t"), c], check=True)
        out = subprocess.run([os.path.join(d, "t")], check=True, capture_output=True, text=True).stdout.split()
    got = [int(x) for x in out]
    exp = [ctypes.sizeof(DgpuConfig), DgpuConfig.lx.offset, DgpuConfig.th_bc_ymin_c1.offset,
           DgpuConfig.pr.offset, DgpuConfig.gy.offset, DgpuConfig.kz.offset]
    assert got == exp


def test_field_doubles_matches_fortran_layout(lib):
    cfg = configs.small_dns(16, 17, 32)
    p = Problem(cfg)
    lib.dgpu_field_doubles.argtypes = [ctypes.c_void_p]
    lib.dgpu_field_doubles.restype = ctypes.c_int64
    n = lib.dgpu_field_doubles(ctypes.cast(p.cconfig.ptr, ctypes.c_void_p))
    assert n == 2 * (cfg.nx // 2 + 1) * (cfg.nz + 2) * (cfg.ny + 2)
    assert n == int(np.prod((cfg.ny + 2, cfg.nz + 2, cfg.nx + 2)))  # = the physical view


def test_create_rejects_bad_sizes_without_touching_the_gpu(lib):
    # validation runs before any HIP call, so this is safe on a CPU-only host
    cfg = configs.small_dns(16, 17, 16)
    p = Problem(cfg)
    s = DgpuConfig.from_buffer_copy(p.cconfig.struct)
    s.nx = 48  # not a power of two
    h = ctypes.c_void_p()
    lib.dgpu_create.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
    lib.dgpu_last_error.restype = ctypes.c_char_p
    rc = lib.dgpu_create(ctypes.cast(ctypes.byref(s), ctypes.c_void_p), ctypes.byref(h))
    assert rc == 2 and b"power" in lib.dgpu_last_error()
    assert not h.value

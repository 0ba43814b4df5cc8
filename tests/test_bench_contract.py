"""bench.py's host-side contract on the CPU: one GPU only (--gpus above 1 exits
with a message), the algorithmic byte model of the DNS kernels at 256x129x256
FP64 (the figures the roofline and the profiles quote, recomputed here from the
field sizes), and the committed traffic summary it reads for the dominant kernel."""
import os
import subprocess
import sys

import bench
from diablo3_amd import configs
from diablo3_amd.config import FP64

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2 and "one GPU only" in r.stderr


def test_dns_256_byte_model():
    cfg = configs.dns_256(FP64)
    kb = bench.kernel_bytes(cfg, False)
    kxr, kzr, ny2, nz = cfg.kxr, cfg.kzr, cfg.ny2, cfg.nz
    C = kxr * kzr * 16 * ny2          # one packed Fourier field
    H = kxr * nz * 16 * ny2           # one x-spectrum field (complex double)
    jlo = min([cfg.jstart, 2] + list(cfg.jstart_th[:cfg.n_th]))  # 1: the scalar's rows start at 1
    nj = cfg.ny - jlo + 1             # planes jlo .. ny carry products
    Hj = kxr * nz * 16 * nj
    assert kb["k_inv_z"] == 4 * C + 4 * H
    # the x pass reads the 4 field spectra of planes jlo-1 .. ny+1 and writes 8 product spectra
    assert kb["k_xmarch"] == 4 * kxr * nz * 16 * (nj + 2) + 8 * Hj == 548110336
    assert kb["k_project"] == 8 * C
    assert set(kb) >= {"k_inv_z", "k_xmarch", "k_fwdz_nl", "k_cn", "k_project"}


def test_traffic_summary_of_the_dominant_kernel():
    t = bench.read_traffic("k_xmarch", "dns_256_fp64")
    assert t is not None and 1.0 <= t / 548110336 <= 1.1

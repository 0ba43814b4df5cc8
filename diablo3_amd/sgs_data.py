"""Training data for the learned closure from box-filtered DNS.

A DNS state of this code (any grid) is filtered onto an LES grid of the same
domain: a top-hat of width fx*dx (fz*dz) centred on each LES point in x and z
(trapezoid weights over the fx+1 DNS points it covers), and in y the average
over each LES cell (u1, u3, th at GYF) or dual cell (u2 at GY), taking the DNS
values as constant over their own cells, so grids that do not nest are handled
exactly.  From the same filter:

* the filtered state (uploaded to an LES handle: its features, at the GYF points
  of rows jstart..Nyp-1, come from dgpu_les_features, the same device function
  the LES step evaluates);
* the subgrid stress tau_ij = bar(u_i u_j) - bar(u_i) bar(u_j) and flux
  q_j = bar(u_j th) - bar(u_j) bar(th) at the LES GYF points, and from them the
  least-squares eddy viscosity and diffusivity
      nu* = max(0, -tau^d_ij S_ij / (2 S_kl S_kl)),  kappa* = max(0, -q_j d_j th / |grad th|^2)
  with S and grad th of the filtered field on the LES grid (spectral in x, z;
  second-order differences in y).

The network's targets are nu*/scale and kappa*/scale with the per-point scale
C^2 D2 sqrt(2 s2) of dgpu_les_features.  Host-side numpy: this prepares data
for the device's training kernels and runs once per DNS snapshot.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, List

import numpy as np

from . import layout
from .config import Config
from .ic import FourierState


def tophat_xz(f: np.ndarray, fx: int, fz: int) -> np.ndarray:
    """[j][z][x] on the DNS grid -> [j][z/fz][x/fx] (periodic, centred top-hat)."""
    def one(a, fac, axis):
        if fac == 1:
            return a
        w = np.ones(fac + 1)
        w[0] = w[-1] = 0.5
        w /= fac
        acc = np.zeros_like(a)
        for m in range(fac + 1):
            acc += w[m] * np.roll(a, -(m - fac // 2), axis=axis)
        sl = [slice(None)] * a.ndim
        sl[axis] = slice(None, None, fac)
        return acc[tuple(sl)]
    return one(one(f, fx, 2), fz, 1)


def overlap_matrix(src_lo, src_hi, dst_lo, dst_hi) -> np.ndarray:
    """M[d][s] = |[dst] ∩ [src]| / |[dst] ∩ union(src)|: the average over each
    destination interval of a function constant on the source intervals."""
    lo = np.maximum(dst_lo[:, None], src_lo[None, :])
    hi = np.minimum(dst_hi[:, None], src_hi[None, :])
    M = np.clip(hi - lo, 0.0, None)
    s = M.sum(axis=1, keepdims=True)
    return np.divide(M, s, out=np.zeros_like(M), where=s > 0)


@dataclass
class YFilter:
    gyf: np.ndarray   # [ny_d + 2 -> ny_l + 2] rows 1..ny_l of GYF fields
    gy: np.ndarray    # rows 2..ny_l of GY fields


def y_filters(dns: Dict[str, np.ndarray], les: Dict[str, np.ndarray]) -> YFilter:
    """Overlap matrices mapping the DNS rows to the LES rows (metrics dicts of grid.metrics)."""
    nd, nl = len(dns["gy"]) - 2, len(les["gy"]) - 2
    # GYF fields: cells [GY(k), GY(k+1)], k = 1..ny
    Mf = np.zeros((nl + 2, nd + 2))
    Mf[1:nl + 1, 1:nd + 1] = overlap_matrix(dns["gy"][1:nd + 1], dns["gy"][2:nd + 2],
                                            les["gy"][1:nl + 1], les["gy"][2:nl + 2])
    # GY fields: dual cells [GYF(k-1), GYF(k)], k = 2..ny
    Mg = np.zeros((nl + 2, nd + 2))
    Mg[2:nl + 1, 2:nd + 1] = overlap_matrix(dns["gyf"][1:nd], dns["gyf"][2:nd + 1],
                                            les["gyf"][1:nl], les["gyf"][2:nl + 1])
    return YFilter(Mf, Mg)


def apply_y(M: np.ndarray, f: np.ndarray) -> np.ndarray:
    return np.tensordot(M, f, axes=(1, 0))


def ddx(f: np.ndarray, length: float, axis: int) -> np.ndarray:
    n = f.shape[axis]
    k = 2.0 * np.pi / length * np.fft.fftfreq(n, 1.0 / n)
    if n % 2 == 0:
        k[n // 2] = 0.0
    sh = [1] * f.ndim
    sh[axis] = n
    return np.real(np.fft.ifft(1j * k.reshape(sh) * np.fft.fft(f, axis=axis), axis=axis))


@dataclass
class Filtered:
    state: FourierState      # the filtered field on the LES grid
    nu_star: np.ndarray      # [rows jstart..ny-1][z][x]
    kappa_star: np.ndarray


# the filtered fields, in the order of dgpu_box_filter (include/dgpu_closure.h):
# u1, u2 at GY, u2 at GYF, u3, th, then the products of (u1, u2 at GYF, u3) and th
BOX_FIELDS = ("u1", "u2g", "u2", "u3", "th", "u1u1", "u1u2", "u1u3", "u2u2", "u2u3", "u3u3",
              "u1th", "u2th", "u3th")


def filter_fields(dns_cfg: Config, dns_met: dict, phys: Dict[str, np.ndarray],
                  les_cfg: Config, les_met: dict) -> np.ndarray:
    """The 14 filtered fields [14][ny_l + 2][nz_l][nx_l] of dgpu_box_filter, in
    numpy (the restatement the device filter is tested against).  phys: u1, u2,
    u3, th as [j][z][x] physical arrays of the DNS (rows 0..ny+1)."""
    fx, fz = dns_cfg.nx // les_cfg.nx, dns_cfg.nz // les_cfg.nz
    assert fx * les_cfg.nx == dns_cfg.nx and fz * les_cfg.nz == dns_cfg.nz and fx % 2 == 0 and fz % 2 == 0
    yf = y_filters(dns_met, les_met)
    nd = dns_cfg.ny
    u1, u2, u3, th = (phys[k] for k in ("u1", "u2", "u3", "th"))
    # u2 at the DNS GYF points (the DNS's own interpolation)
    u2f = np.zeros_like(u2)
    u2f[1:nd + 1] = 0.5 * (u2[1:nd + 1] + u2[2:nd + 2])

    def F(a, gy=False):
        return apply_y(yf.gy if gy else yf.gyf, tophat_xz(a, fx, fz))

    return np.stack([F(u1), F(u2, gy=True), F(u2f), F(u3), F(th),
                     F(u1 * u1), F(u1 * u2f), F(u1 * u3), F(u2f * u2f), F(u2f * u3), F(u3 * u3),
                     F(u1 * th), F(u2f * th), F(u3 * th)])


def y_weights(dns_met: dict, les_met: dict):
    """(mf, mg): the row weights dgpu_box_filter takes, C-contiguous."""
    yf = y_filters(dns_met, les_met)
    return np.ascontiguousarray(yf.gyf), np.ascontiguousarray(yf.gy)


def targets(les_cfg: Config, les_met: dict, fields: np.ndarray) -> "Filtered":
    """From the 14 filtered fields: the filtered state, the subgrid stress and flux,
    and the least-squares nu*, kappa* on rows jstart..ny-1 of the LES grid."""
    d = dict(zip(BOX_FIELDS, fields))
    b1, b2g, b2, b3, bt = d["u1"], d["u2g"], d["u2"], d["u3"], d["th"]
    # GYF-located filtered velocities, for the stresses and the strain
    vel = [b1, b2, b3]
    names = ("u1", "u2", "u3")
    tau = {}
    for i in range(3):
        for k in range(i, 3):
            tau[i, k] = d[names[i] + names[k]] - vel[i] * vel[k]
    q = [d[names[i] + "th"] - vel[i] * bt for i in range(3)]
    # strain and scalar gradient of the filtered field on the LES grid
    ny_l = les_cfg.ny
    rows = slice(les_cfg.jstart, ny_l)
    y = les_met["gyf"]

    def grad(a):
        gx = ddx(a, les_cfg.lx, 2)
        gz = ddx(a, les_cfg.lz, 1)
        gyv = np.zeros_like(a)
        gyv[1:ny_l + 1] = np.gradient(a[1:ny_l + 1], y[1:ny_l + 1], axis=0)
        return [gx, gyv, gz]

    G = [grad(v) for v in vel]           # G[i][k] = d u_i / d x_k
    S = {(i, k): 0.5 * (G[i][k] + G[k][i]) for i in range(3) for k in range(3)}
    tr = (tau[0, 0] + tau[1, 1] + tau[2, 2]) / 3.0
    num = np.zeros_like(b1)
    den = np.zeros_like(b1)
    for i in range(3):
        for k in range(3):
            t = tau[min(i, k), max(i, k)] - (tr if i == k else 0.0)
            num += t * S[i, k]
            den += S[i, k] * S[i, k]
    nu = np.maximum(0.0, -num / np.where(den > 0, 2.0 * den, np.inf))
    gt = grad(bt)
    gn = gt[0] ** 2 + gt[1] ** 2 + gt[2] ** 2
    kq = q[0] * gt[0] + q[1] * gt[1] + q[2] * gt[2]
    ka = np.maximum(0.0, -kq / np.where(gn > 0, gn, np.inf))
    # the filtered state: GYF fields on rows 1..ny, u2 on rows 2..ny (walls from the BCs)
    st = FourierState(
        cu1=layout.fft_xz_to_fourier(les_cfg, b1), cu2=layout.fft_xz_to_fourier(les_cfg, b2g),
        cu3=layout.fft_xz_to_fourier(les_cfg, b3), cp=layout.compact_zeros(les_cfg),
        cth=[layout.fft_xz_to_fourier(les_cfg, bt)])
    return Filtered(st, nu[rows], ka[rows])


def filter_snapshot(dns_cfg: Config, dns_met: dict, phys: Dict[str, np.ndarray],
                    les_cfg: Config, les_met: dict) -> Filtered:
    """phys: u1, u2, u3, th as [j][z][x] physical arrays of the DNS (rows 0..ny+1)."""
    return targets(les_cfg, les_met, filter_fields(dns_cfg, dns_met, phys, les_cfg, les_met))


def make_samples(X: np.ndarray, scale: np.ndarray, inside: np.ndarray, nu: np.ndarray, ka: np.ndarray,
                 cap_quantile: float = 0.995):
    """Training pairs from features and targets of one snapshot: the points with
    a positive scale, targets nu*/scale, kappa*/scale capped at their quantile."""
    nu, ka = nu.ravel(), ka.ravel()
    ok = (scale > 0) & np.isfinite(nu) & np.isfinite(ka)
    T = np.stack([nu[ok] / scale[ok], ka[ok] / scale[ok]], axis=1)
    for c in range(2):
        cap = np.quantile(T[:, c], cap_quantile)
        T[:, c] = np.minimum(T[:, c], cap)
    return X[ok], T.astype(np.float32), ok

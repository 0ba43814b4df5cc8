"""Run configuration: the reference's input decks, flattened into ``dgpu_config``.

The reference reads three positional text decks (``input.dat``, ``input_chan.dat``,
``grid_def.all``) into module globals (for/parameters.f90:75-291,
for/make_headers:11-30).  :class:`Config` holds the same values, :func:`read_decks`
parses the decks with the reference's line structure, and :meth:`Config.derive`
applies the derived settings the reference computes at start-up:

* ``nu = 1/Re`` (for/parameters.f90:102)
* ``Ro_inv = 1/ro`` (for/parameters.f90:224)
* the infinite/finite-front background gradient and Neumann-BC compensation
  (for/parameters.f90:257-277)
* ``jstart/jend`` from init_chan_mpi at one rank (for/domain.f90:318-347).  With a
  single rank the ``rankY == 0`` branch is taken, so ``jend = Nyp`` whatever the
  top BC is; the top BC rows then override row Nyp in every implicit solve.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

MAX_TH = 4
DIRICHLET, NEUMANN = 0, 1
FP64, MIXED = 0, 1


@dataclass
class Config:
    # grid_def.all (code axes: y is wall-normal)
    nx: int = 64
    ny: int = 65
    nz: int = 64
    n_th: int = 1
    # input.dat
    flavor: str = "Front"
    use_les: bool = False
    re: float = 1000.0
    beta: float = 1.0
    lx: float = 2.0 * math.pi
    ly: float = 2.0
    lz: float = 2.0 * math.pi
    nu_v_scale: float = 1.0
    create_new_flow: bool = True
    delta_t: float = 1e-3
    variable_dt: bool = False
    create_new_th: List[bool] = field(default_factory=lambda: [True])
    ri: List[float] = field(default_factory=lambda: [1.0])
    pr: List[float] = field(default_factory=lambda: [1.0])
    # input_chan.dat
    time_ad_meth: int = 1
    les_model_type: int = 5
    ic_type: int = 4
    kick: float = 2e-3
    physical_noise: bool = True
    ro: float = 1.0
    delta: float = 1.0
    grav_x: float = 0.0
    grav_y: float = 1.0
    grav_z: float = 0.0
    f_type: int = 3
    ubulk0: float = 0.0
    px0: float = 0.0
    omega0: float = 0.0
    amp_omega0: float = 0.0
    force_start: float = 0.0
    u_bc_ymin: int = DIRICHLET
    u_bc_ymin_c1: float = 0.0
    w_bc_ymin: int = DIRICHLET
    w_bc_ymin_c1: float = 0.0
    v_bc_ymin: int = DIRICHLET
    v_bc_ymin_c1: float = 0.0
    u_bc_ymax: int = NEUMANN
    u_bc_ymax_c1: float = 0.0
    w_bc_ymax: int = NEUMANN
    w_bc_ymax_c1: float = 0.0
    v_bc_ymax: int = DIRICHLET
    v_bc_ymax_c1: float = 0.0
    th_bc_ymin: List[int] = field(default_factory=lambda: [NEUMANN])
    th_bc_ymin_c1: List[float] = field(default_factory=lambda: [1.0])
    th_bc_ymax: List[int] = field(default_factory=lambda: [NEUMANN])
    th_bc_ymax_c1: List[float] = field(default_factory=lambda: [1.0])
    # grid generation (proc/init/create_grid.m)
    grid_type: int = 1
    grid_cs: float = 1.75
    # explicit background scalar gradients; None = derive from ic_type as the reference
    dthdx: Optional[List[float]] = None
    dthdz: Optional[List[float]] = None
    # apply the IC-type Neumann compensation of input_chan (False when the BC values
    # already hold it)
    bc_compensation: bool = True
    # the library's own switches
    precision: int = FP64
    # derived (filled by derive())
    nu: float = 0.0
    ro_inv: float = 0.0
    jstart: int = 0
    jend: int = 0
    jstart_th: List[int] = field(default_factory=list)
    jend_th: List[int] = field(default_factory=list)

    # ---- sizes (for/domain.f90:20-32, one rank) ----
    @property
    def nxp(self) -> int:
        return self.nx // 2

    @property
    def nkx(self) -> int:
        return self.nx // 3

    @property
    def nkz(self) -> int:
        return self.nz // 3

    @property
    def twonkz(self) -> int:
        return 2 * self.nkz

    @property
    def kxr(self) -> int:
        return self.nkx + 1

    @property
    def kzr(self) -> int:
        return self.twonkz + 1

    @property
    def ny2(self) -> int:
        return self.ny + 2

    def copy(self, **kw) -> "Config":
        c = dataclasses.replace(self, **kw)
        for name in ("create_new_th", "ri", "pr", "th_bc_ymin", "th_bc_ymin_c1",
                     "th_bc_ymax", "th_bc_ymax_c1", "jstart_th", "jend_th"):
            setattr(c, name, list(getattr(c, name)))
        return c

    def derive(self) -> "Config":
        """Apply the start-up derivations of init_parameters/input_chan/init_chan_mpi."""
        c = self.copy()
        c.nu = 1.0 / c.re                      # for/parameters.f90:102
        c.ro_inv = 1.0 / c.ro                  # for/parameters.f90:224
        nth = c.n_th
        for name in ("ri", "pr", "th_bc_ymin", "th_bc_ymin_c1", "th_bc_ymax",
                     "th_bc_ymax_c1", "create_new_th"):
            v = getattr(c, name)
            if len(v) < nth:
                raise ValueError(f"{name} needs {nth} entries")
        dthdx = [0.0] * nth
        dthdz = [0.0] * nth
        # for/parameters.f90:257-277
        comp = c.bc_compensation
        if c.ic_type in (4, 5):
            if comp and c.w_bc_ymin == NEUMANN:
                c.w_bc_ymin_c1 -= 1.0
            if comp and c.w_bc_ymax == NEUMANN:
                c.w_bc_ymax_c1 -= 1.0
            if nth:
                dthdx[0] = c.ro_inv / c.delta
        elif c.ic_type in (6, 7, 8):
            if comp and c.w_bc_ymin == NEUMANN:
                c.w_bc_ymin_c1 -= 2.0 * c.delta / c.lx
            if comp and c.w_bc_ymax == NEUMANN:
                c.w_bc_ymax_c1 -= 2.0 * c.delta / c.lx
            if nth:
                dthdx[0] = 2.0 / c.lx * c.ro_inv
        if c.dthdx is not None:
            dthdx = list(c.dthdx)
        if c.dthdz is not None:
            dthdz = list(c.dthdz)
        c.dthdx, c.dthdz = dthdx, dthdz
        # init_chan_mpi, rankY == 0 branch (for/domain.f90:318-330)
        c.jend = c.ny
        c.jstart = 1 if c.u_bc_ymin == NEUMANN else 2
        c.jend_th = [c.ny] * nth
        c.jstart_th = [1 if c.th_bc_ymin[n] == NEUMANN else 2 for n in range(nth)]
        return c

    def validate(self) -> None:
        if self.nx % 2 or self.nz % 2:
            raise ValueError("nx and nz must be even")
        for n in (self.nx, self.nz):
            if n & (n - 1):
                raise ValueError("nx and nz must be powers of two")
        if self.ny < 4:
            raise ValueError("ny must be >= 4")
        if not 0 <= self.n_th <= MAX_TH:
            raise ValueError("n_th must be 0..4")
        if self.time_ad_meth != 1:
            raise ValueError("only TIME_AD_METH = 1 exists (for/solver.f90:1276-1297)")


def _f(x):
    return float(str(x).replace("D", "E").replace("d", "e"))


def _b(x):
    return str(x).strip().upper().strip(".") in ("TRUE", "T")


def _tok(line: str) -> List[str]:
    return line.replace(",", " ").split()


def read_decks(input_dat: str, input_chan_dat: str, grid_def: str) -> Config:
    """Parse the three decks with the reference's positional layout.

    input.dat: for/parameters.f90:84-121; input_chan.dat: for/parameters.f90:202-250;
    grid_def.all: for/make_headers:11-30 (user Ny/Nz are swapped into code Nz/Ny).
    """
    g = {}
    for line in open(grid_def):
        line = line.split("#")[0].strip()
        if "=" in line:
            k, v = line.split("=")
            g[k.strip().lower()] = int(v)
    c = Config(nx=g["nx"], ny=g["nz"], nz=g["ny"], n_th=g.get("nth", 1))
    lines = open(input_dat).read().splitlines()
    it = iter(lines)

    def nxt():
        return _tok(next(it))

    for _ in range(4):
        next(it)
    t = nxt()
    c.flavor = t[0].strip("'\"")
    if abs(_f(t[1]) - 3.4) > 1e-6:
        raise ValueError("Wrong input data format.")
    next(it)
    t = nxt()
    if not _b(t[0]):
        raise ValueError("Serial processing has been deprecated in diablo3.")
    c.use_les = _b(t[1])
    next(it)
    t = nxt()
    c.re, c.beta, c.lx, c.lz, c.ly = (_f(x) for x in t[:5])
    next(it)
    c.nu_v_scale = _f(nxt()[0])
    next(it)
    t = nxt()
    if int(t[0]) != 2:
        raise ValueError("only the channel (NUM_PER_DIR = 2) exists")
    c.create_new_flow = _b(t[1])
    next(it)
    t = nxt()
    c.delta_t = _f(t[2])
    c.variable_dt = _b(t[4])
    next(it)
    nxt()
    next(it)
    c.create_new_th, c.ri, c.pr = [], [], []
    for _ in range(c.n_th):
        next(it)
        c.create_new_th.append(_b(nxt()[0]))
        next(it)
        nxt()
        next(it)
        t = nxt()
        c.ri.append(_f(t[0]))
        c.pr.append(_f(t[1]))

    lines = open(input_chan_dat).read().splitlines()
    it = iter(lines)
    for _ in range(4):
        next(it)
    if abs(_f(nxt()[0]) - 3.4) > 1e-6:
        raise ValueError("Wrong input data format input_chan")
    next(it)
    c.time_ad_meth = int(nxt()[0])
    next(it)
    c.les_model_type = int(nxt()[0])
    next(it)
    t = nxt()
    c.ic_type, c.kick, c.physical_noise = int(t[0]), _f(t[1]), _b(t[2])
    next(it)
    c.ro = _f(nxt()[0])
    next(it)
    c.delta = _f(nxt()[0])
    next(it)
    t = nxt()
    c.grav_x, c.grav_z, c.grav_y = (_f(x) for x in t[:3])
    next(it)
    t = nxt()
    c.f_type = int(t[0])
    c.ubulk0, c.px0, c.omega0, c.amp_omega0 = (_f(x) for x in t[1:5])
    c.force_start = _f(t[5]) if len(t) > 5 else 0.0
    next(it)
    next(it)
    pairs = []
    for _ in range(6):
        t = nxt()
        pairs.append((int(t[0]), _f(t[1])))
        next(it)
    (c.u_bc_ymin, c.u_bc_ymin_c1), (c.w_bc_ymin, c.w_bc_ymin_c1), \
        (c.v_bc_ymin, c.v_bc_ymin_c1), (c.u_bc_ymax, c.u_bc_ymax_c1), \
        (c.w_bc_ymax, c.w_bc_ymax_c1), (c.v_bc_ymax, c.v_bc_ymax_c1) = pairs
    c.th_bc_ymin, c.th_bc_ymin_c1, c.th_bc_ymax, c.th_bc_ymax_c1 = [], [], [], []
    for _ in range(c.n_th):
        next(it)
        t = nxt()
        c.th_bc_ymin.append(int(t[0]))
        c.th_bc_ymin_c1.append(_f(t[1]))
        next(it)
        t = nxt()
        c.th_bc_ymax.append(int(t[0]))
        c.th_bc_ymax_c1.append(_f(t[1]))
    return c


# --------------------------------------------------------------------------------------
# ctypes mirror of dgpu_config (include/diablo_gpu.h)
# --------------------------------------------------------------------------------------
_I4 = ctypes.c_int32 * MAX_TH
_D4 = ctypes.c_double * MAX_TH
_PD = ctypes.POINTER(ctypes.c_double)


class DgpuConfig(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("n_th", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("use_les", ctypes.c_int32), ("les_model_type", ctypes.c_int32),
        ("flavor_front", ctypes.c_int32), ("f_type", ctypes.c_int32),
        ("jstart", ctypes.c_int32), ("jend", ctypes.c_int32),
        ("jstart_th", _I4), ("jend_th", _I4),
        ("u_bc_ymin", ctypes.c_int32), ("v_bc_ymin", ctypes.c_int32),
        ("w_bc_ymin", ctypes.c_int32), ("u_bc_ymax", ctypes.c_int32),
        ("v_bc_ymax", ctypes.c_int32), ("w_bc_ymax", ctypes.c_int32),
        ("th_bc_ymin", _I4), ("th_bc_ymax", _I4), ("homogeneous_x", ctypes.c_int32),
        ("reserved_i", ctypes.c_int32 * 3),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("nu", ctypes.c_double), ("beta", ctypes.c_double), ("nu_v_scale", ctypes.c_double),
        ("ro_inv", ctypes.c_double), ("delta", ctypes.c_double),
        ("grav_x", ctypes.c_double), ("grav_y", ctypes.c_double), ("grav_z", ctypes.c_double),
        ("px0", ctypes.c_double), ("omega0", ctypes.c_double),
        ("amp_omega0", ctypes.c_double), ("force_start", ctypes.c_double),
        ("u_bc_ymin_c1", ctypes.c_double), ("v_bc_ymin_c1", ctypes.c_double),
        ("w_bc_ymin_c1", ctypes.c_double), ("u_bc_ymax_c1", ctypes.c_double),
        ("v_bc_ymax_c1", ctypes.c_double), ("w_bc_ymax_c1", ctypes.c_double),
        ("th_bc_ymin_c1", _D4), ("th_bc_ymax_c1", _D4),
        ("pr", _D4), ("dthdx", _D4), ("dthdz", _D4), ("reserved_d", _D4),
        ("gy", _PD), ("gyf", _PD), ("dy", _PD), ("dyf", _PD), ("kx", _PD), ("kz", _PD),
    ]


class CConfig:
    """A DgpuConfig plus the numpy arrays its pointers refer to (kept alive here)."""

    def __init__(self, cfg: Config, metrics: dict, kx: np.ndarray, kz: np.ndarray):
        if cfg.nu == 0.0:
            raise ValueError("call Config.derive() first")
        s = DgpuConfig()
        for name in ("nx", "ny", "nz", "n_th", "precision", "les_model_type", "f_type",
                     "jstart", "jend", "u_bc_ymin", "v_bc_ymin", "w_bc_ymin", "u_bc_ymax",
                     "v_bc_ymax", "w_bc_ymax"):
            setattr(s, name, int(getattr(cfg, name)))
        s.use_les = int(bool(cfg.use_les))
        s.flavor_front = int(cfg.flavor.strip() == "Front")
        # homogeneousX (for/parameters.f90:282-287)
        s.homogeneous_x = int(cfg.ic_type not in (5, 6, 7, 8))
        for name in ("lx", "ly", "lz", "nu", "beta", "nu_v_scale", "ro_inv", "delta",
                     "grav_x", "grav_y", "grav_z", "px0", "omega0", "amp_omega0",
                     "force_start", "u_bc_ymin_c1", "v_bc_ymin_c1", "w_bc_ymin_c1",
                     "u_bc_ymax_c1", "v_bc_ymax_c1", "w_bc_ymax_c1"):
            setattr(s, name, float(getattr(cfg, name)))
        for n in range(cfg.n_th):
            s.jstart_th[n] = cfg.jstart_th[n]
            s.jend_th[n] = cfg.jend_th[n]
            s.th_bc_ymin[n] = cfg.th_bc_ymin[n]
            s.th_bc_ymax[n] = cfg.th_bc_ymax[n]
            s.th_bc_ymin_c1[n] = cfg.th_bc_ymin_c1[n]
            s.th_bc_ymax_c1[n] = cfg.th_bc_ymax_c1[n]
            s.pr[n] = cfg.pr[n]
            s.dthdx[n] = cfg.dthdx[n]
            s.dthdz[n] = cfg.dthdz[n]
        self._keep = {k: np.ascontiguousarray(metrics[k], dtype=np.float64)
                      for k in ("gy", "gyf", "dy", "dyf")}
        self._keep["kx"] = np.ascontiguousarray(kx, dtype=np.float64)
        self._keep["kz"] = np.ascontiguousarray(kz, dtype=np.float64)
        for k, a in self._keep.items():
            setattr(s, k, a.ctypes.data_as(_PD))
        self.struct = s
        self.cfg = cfg

    @property
    def ptr(self):
        return ctypes.byref(self.struct)

"""Everything the host derives from a Config before the first step.

This is the host half of init_parameters + init_flow (for/parameters.f90:75-188,
for/flow.f90:100-170): grid file -> metrics, wavenumbers, the flattened C config,
and the initial Fourier state.
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from . import grid, ic
from .config import CConfig, Config


class Problem:
    def __init__(self, cfg: Config, gy_file: Optional[np.ndarray] = None):
        cfg.validate()
        if cfg.nu == 0.0:
            cfg = cfg.derive()
        self.cfg = cfg
        self.gy_file = (grid.create_grid(cfg.ny, cfg.ly, cfg.grid_cs, cfg.grid_type)
                        if gy_file is None else np.asarray(gy_file, dtype=np.float64))
        if len(self.gy_file) != cfg.ny + 1:
            raise ValueError("Grid file and program do not match.")  # for/domain.f90:498-506
        self.met = grid.metrics(self.gy_file)
        self.kx, self.kz = grid.wavenumbers(cfg.nx, cfg.nz, cfg.lx, cfg.lz)
        self.cconfig = CConfig(cfg, self.met, self.kx, self.kz)

    def initial_state(self, seed: int = 0, theta_noise: float = 0.0) -> ic.FourierState:
        """IC before projection; project with Solver.rem_div (for/flow.f90:487)."""
        return ic.initial_state(self.cfg, self.met, seed, theta_noise)

"""diablo3_amd — MI355X-native RKW3 + Crank–Nicolson time step of DIABLO's channel solver.

The hot path is ``libdiablo_gpu.so`` (hand-written HIP for gfx950 behind the C ABI in
``include/diablo_gpu.h``).  This package is its host side: the input decks, grid,
wavenumbers, initial-condition hooks, Fortran-layout helpers and a ctypes binding.
"""
from .config import Config, read_decks, FP64, MIXED, DIRICHLET, NEUMANN  # noqa: F401
from . import grid, layout, ic, configs  # noqa: F401

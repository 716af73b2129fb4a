"""anechoic_amd: MI355X-native (gfx950) SCA annealing engine with Onsager echo correction.

The anneals run in hand-written HIP kernels (libanechoic_mi355x.so, C ABI in include/anechoic.h) reached through
ctypes; torch provides device memory and streams.
"""
from .engine import Engine, np_of, run, spins_pm, tables  # noqa: F401

__all__ = ["Engine", "run", "tables", "np_of", "spins_pm"]

"""MI355X-native Truncated Wigner Monte Carlo trajectory engine.

Hand-written HIP kernels for gfx950 behind the C ABI in include/twmc_hip.h
(libtwmc_hip.so), a C++ host layer that re-implements the reference's
Solver / Task / Manager plugin boundary (pincosimulator_amd/host), and this
Python front end used by the tests and the benchmark.
"""
from .engine import Engine, Result, Sweep, System, frame_schedule, noisy_matrix  # noqa: F401
from . import observables  # noqa: F401
from ._lib import TWMCError, LIB_PATH  # noqa: F401

__all__ = ["Engine", "Result", "Sweep", "System", "frame_schedule", "noisy_matrix", "observables",
           "TWMCError", "LIB_PATH"]

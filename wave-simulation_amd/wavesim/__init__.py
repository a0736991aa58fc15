"""wavesim: MI355X-native forward-solver time loop of WAVE-Simulation.

The compute lives in libwavehip.so (hand-written gfx950 HIP kernels behind the
C ABI in include/wave_abi.h); this package is the thin host-side binding.
"""
from ._abi import WaveError, load  # noqa: F401
from .solver import ForwardSolver, create  # noqa: F401
from .acquisition import ricker, source_signal  # noqa: F401

"""qbg: MI355X-native Hamiltonian application for the Qbox plane-wave path.

The compute lives in libqbg.so (hand-written HIP kernels for gfx950 behind
the C ABI of include/qbg.h); this package is the Python harness over it.
"""
from . import configs  # noqa: F401
from ._lib import (Context, DArray, EnergyFunctional, Graph, Grid, Mixer,  # noqa: F401
                   NonLocalPotential, Profiler, QbgError, WfStepper, ENERGY_NAMES, fft_splits,
                   gram, gram_complex, gram_noalloc, load, polar, stream_join)

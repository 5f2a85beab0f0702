"""hmclab_amd — MI355X-native HMC engine for hmclab's hot path.

The package mirrors the reference's user-facing names for the HMC path
(``Distributions``, ``MassMatrices``, ``Samplers.HMC``, ``Samples``) and runs
the trajectories, gradients, Hamiltonians and Metropolis tests in
hand-written gfx950 HIP kernels (``libhmc.so``, C-ABI in
``include/hmc_engine.h``).
"""
from . import Distributions, MassMatrices, Samplers, Samples, diagnostics, tomography  # noqa: F401
from ._lib import device_count  # noqa: F401
from .Samplers import HMC, ParallelSampleSMP  # noqa: F401

__version__ = "0.1.0"

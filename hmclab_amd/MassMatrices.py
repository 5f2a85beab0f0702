"""Mass matrices with the reference's interface (hmclab/MassMatrices.py:27-238).

``Unit`` and ``Diagonal`` describe the metric; inside sampling the kinetic
energy, its gradient and the momentum draw run on the GPU (fused into the
trajectory kernels).  The host methods below keep the reference's signatures
and arithmetic for direct use and for host-generated (reference-ordered)
momenta.
"""
from __future__ import annotations

import numpy as np


class _AbstractMassMatrix:
    """MassMatrices.py:27-79."""

    name: str = "mass matrix abstract base class"
    dimensions: int = -1
    rng: np.random.Generator = np.random.default_rng()

    def kinetic_energy(self, momentum):
        raise NotImplementedError

    def kinetic_energy_gradient(self, momentum, position=None, g=None):
        raise NotImplementedError

    def generate_momentum(self):
        raise NotImplementedError

    def accept(self):
        pass

    def reject(self):
        pass

    # device description: None for unit mass, else the diagonal (d,)
    def _device_diagonal(self):
        raise NotImplementedError


class Unit(_AbstractMassMatrix):
    """Identity metric (MassMatrices.py:82-156)."""

    def __init__(self, dimensions: int = -1, rng: np.random.Generator = None):
        self.name = "unit mass matrix"
        self.dimensions = dimensions
        if rng is not None:
            self.rng = rng

    def kinetic_energy(self, momentum: np.ndarray) -> float:
        if momentum.shape != (self.dimensions, 1):
            raise ValueError()
        return 0.5 * (momentum.T @ momentum).item(0)

    def kinetic_energy_gradient(self, momentum, position=None, g=None) -> np.ndarray:
        if momentum.shape != (self.dimensions, 1):
            raise ValueError()
        return momentum

    def generate_momentum(self, rng: np.random.Generator = None) -> np.ndarray:
        return (rng or self.rng).normal(size=(self.dimensions, 1))

    @property
    def matrix(self) -> np.ndarray:
        return np.eye(self.dimensions)

    def _device_diagonal(self):
        return None

    @staticmethod
    def create_default(dimensions: int, rng: np.random.Generator = None) -> "Unit":
        return Unit(dimensions, rng)


class Diagonal(_AbstractMassMatrix):
    """Diagonal metric (MassMatrices.py:159-238)."""

    def __init__(self, diagonal: np.ndarray, rng: np.random.Generator = None):
        self.name = "diagonal mass matrix"
        diagonal = np.array(diagonal, dtype=np.float64)
        diagonal.shape = (diagonal.size, 1)
        self.dimensions = diagonal.size
        self.diagonal = diagonal
        self.inverse_diagonal = 1.0 / self.diagonal
        if rng is not None:
            self.rng = rng

    def kinetic_energy(self, momentum: np.ndarray) -> float:
        if momentum.shape != (self.dimensions, 1):
            raise ValueError()
        return 0.5 * np.vdot(momentum, self.inverse_diagonal * momentum)

    def kinetic_energy_gradient(self, momentum, position=None, g=None) -> np.ndarray:
        if momentum.shape != (self.dimensions, 1):
            raise ValueError()
        return self.inverse_diagonal * momentum

    def generate_momentum(self, rng: np.random.Generator = None) -> np.ndarray:
        return np.sqrt(self.diagonal) * (rng or self.rng).normal(size=(self.dimensions, 1))

    @property
    def matrix(self) -> np.ndarray:
        return np.diagflat(self.diagonal)

    def _device_diagonal(self):
        return self.diagonal.reshape(-1)

    @staticmethod
    def create_default(dimensions: int, rng: np.random.Generator = None) -> "Diagonal":
        return Diagonal(np.ones((dimensions, 1)), rng=rng)

"""Gauss-Hermite nodes (src/smolyax/nodes/gausshermite.py).

Node values are mean + scaling * (roots of H_{n+1}) from numpy.polynomial.hermite.hermgauss, and
quadrature weights are hermgauss weights / sqrt(pi), i.e. the probability measure
exp(-x^2)/sqrt(pi) dx (gausshermite.py:57-62, 81-87).  The sequence is not nested.
"""

from functools import lru_cache

import numpy as np

from .base import Generator, Generator1D


@lru_cache(maxsize=None)
def _hermgauss(n: int):
    x, w = np.polynomial.hermite.hermgauss(n + 1)
    return x, w / np.sqrt(np.pi)


class GaussHermite1D(Generator1D):
    """gausshermite.py:10-164."""

    def __init__(self, mean: float = 0.0, scaling: float = 1.0) -> None:
        super().__init__(is_nested=False)
        self._mean = mean
        self._scaling = scaling
        self._cache = {}

    @property
    def mean(self) -> float:
        return self._mean

    @property
    def scaling(self) -> float:
        return self._scaling

    def __call__(self, n: int) -> np.ndarray:
        v = self._cache.get(n)
        if v is None:
            v = self.scale(_hermgauss(n)[0])
            self._cache[n] = v
        return v

    def get_quadrature_weights(self, n: int) -> np.ndarray:
        return _hermgauss(n)[1]

    def scale(self, x):
        return self._mean + self._scaling * x

    def scale_back(self, x):
        return (x - self._mean) / self._scaling

    def get_random(self, n: int = 1) -> np.ndarray:
        """Samples of N(mean, scaling^2 / 2) (gausshermite.py:126-145)."""
        return self.scale(np.random.randn(n) / np.sqrt(2))

    def _table_key(self):
        return ("gh", float(self._mean), float(self._scaling))

    def __repr__(self) -> str:
        return f"Gauss-Hermite (mean = {self._mean}, scaling = {self._scaling})"


class GaussHermite(Generator):
    """gausshermite.py:167-249."""

    def __init__(self, mean=None, scaling=None, dim: int = None):
        if dim is None:
            if scaling is not None:
                dim = len(scaling)
            elif mean is not None:
                dim = len(mean)
            else:
                raise ValueError("Must specify at least one of 'dim', 'mean', or 'scaling'.")
        if mean is None and scaling is None:
            Generator.__init__(self, [GaussHermite1D()] * dim)
            self._mean = np.zeros(dim)
            self._scaling = np.ones(dim)
        else:
            self._mean = np.zeros(dim) if mean is None else np.asarray(mean, dtype=np.float64)
            self._scaling = np.ones(dim) if scaling is None else np.asarray(scaling, dtype=np.float64)
            Generator.__init__(self, [GaussHermite1D(m, a) for m, a in zip(self._mean, self._scaling)])

    def scale(self, x):
        return self._mean + self._scaling * x

    def scale_back(self, x):
        return (x - self._mean) / self._scaling

    def __repr__(self) -> str:
        return (f"Gauss Hermite (d = {self.dim}, mean = {self._mean.tolist()}, "
                f"scaling = {self._scaling.tolist()})")

"""Node-sequence generators: the abstract interface (src/smolyax/nodes/base.py:8-262).

Host-side setup code.  A `Generator` is a list of per-dimension `Generator1D` objects; the
interpolator asks it for node values, quadrature weights and the scaling between the reference
domain and the user domain.
"""

from abc import ABC, abstractmethod
from typing import Iterator, List

import numpy as np


class Generator1D(ABC):
    """One-dimensional node sequence (base.py:8-128)."""

    def __init__(self, is_nested: bool) -> None:
        self._is_nested = bool(is_nested)

    @property
    def is_nested(self) -> bool:
        return self._is_nested

    @abstractmethod
    def __call__(self, n: int) -> np.ndarray:
        """The first n+1 nodes (node indices 0..n), mapped to the user domain."""

    @abstractmethod
    def scale(self, x):
        """Reference domain -> user domain."""

    @abstractmethod
    def scale_back(self, x):
        """User domain -> reference domain."""

    @abstractmethod
    def get_random(self, n: int = 1) -> np.ndarray:
        """Samples of the measure the quadrature integrates against."""

    @abstractmethod
    def get_quadrature_weights(self, n: int) -> np.ndarray:
        """Weights of the first n+1 nodes."""

    # Key identifying the 1D sequence (used to share tables between dims with equal sequences).
    def _table_key(self):
        return (type(self).__name__, id(self))


class Generator:
    """Per-dimension container (base.py:131-262)."""

    def __init__(self, node_gens: List[Generator1D]):
        assert all(g.is_nested == node_gens[0].is_nested for g in node_gens)
        self._gens = list(node_gens)
        self._is_nested = node_gens[0].is_nested

    @property
    def dim(self) -> int:
        return len(self._gens)

    @property
    def is_nested(self) -> bool:
        return self._is_nested

    def __getitem__(self, i: int) -> Generator1D:
        return self._gens[i]

    def __iter__(self) -> Iterator[Generator1D]:
        for i in range(self.dim):
            yield self._gens[i]

    def __len__(self) -> int:
        return self.dim

    def get_random(self, n: int = 0) -> np.ndarray:
        """n == 0: one point of shape (dim,); otherwise (n, dim) (base.py:201-218)."""
        if n == 0:
            return np.squeeze([g.get_random() for g in self._gens])
        return np.array([g.get_random(n) for g in self._gens]).T

    def scale(self, x):
        x = np.asarray(x)
        assert x.shape[-1] == self.dim
        cols = [g.scale(x[..., i]) for i, g in enumerate(self._gens)]
        return np.stack(cols, axis=-1)

    def scale_back(self, x):
        x = np.asarray(x)
        assert x.shape[-1] == self.dim
        cols = [g.scale_back(x[..., i]) for i, g in enumerate(self._gens)]
        return np.stack(cols, axis=-1)

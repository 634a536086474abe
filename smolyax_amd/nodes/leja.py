"""Leja nodes on [-1, 1] with optional affine domains (src/smolyax/nodes/leja.py).

The sequence is the symmetric Leja-type sequence of the reference: 0, 1, -1, 1/sqrt(2),
-1/sqrt(2), and then for j >= 5 node j is -node(j-1) for even j and sqrt((node((j+1)/2) + 1) / 2)
for odd j, i.e. cosines of bisected angles (leja.py:17,48-57).  The values are produced by the same
IEEE operations so the node table is bit-identical.  Quadrature weights integrate the first n+1
monomials exactly against dx/2 on [-1, 1] (leja.py:83-93).
"""

from functools import lru_cache
from typing import Sequence

import numpy as np

from .base import Generator, Generator1D

_SEQ = [0.0, 1.0, -1.0, 1 / np.sqrt(2), -1 / np.sqrt(2)]


def leja_sequence(n: int) -> np.ndarray:
    """Nodes 0..n on the reference domain."""
    while len(_SEQ) <= n:
        j = len(_SEQ)
        if j % 2 == 0:
            _SEQ.append(-_SEQ[j - 1])
        else:
            _SEQ.append(float(np.sqrt((_SEQ[(j + 1) // 2] + 1) / 2)))
    return np.array(_SEQ[: n + 1], dtype=np.float64)


@lru_cache(maxsize=None)
def _leja_quadrature(n: int) -> np.ndarray:
    q = leja_sequence(n)
    vander = np.vstack([q**i for i in range(n + 1)])
    moments = np.array([(1 + (-1) ** i) / (2.0 * (i + 1)) for i in range(n + 1)])
    return np.linalg.solve(vander, moments)


def _affine(x, src, dst):
    x = (x - src[0]) / (src[1] - src[0])
    return x * (dst[1] - dst[0]) + dst[0]


class Leja1D(Generator1D):
    """leja.py:10-204."""

    def __init__(self, domain=None) -> None:
        super().__init__(is_nested=True)
        self._domain = None if domain is None else np.asarray(domain, dtype=np.float64)
        self._ref = None if domain is None else (-1, 1)
        self._cache = {}

    @property
    def domain(self):
        return self._domain

    def __call__(self, n: int) -> np.ndarray:
        v = self._cache.get(n)
        if v is None:
            v = self.scale(leja_sequence(n))
            self._cache[n] = v
        return v

    def get_quadrature_weights(self, n: int) -> np.ndarray:
        return _leja_quadrature(n)

    def scale(self, x, d1=None, d2=None):
        d1 = self._ref if d1 is None else d1
        d2 = self._domain if d2 is None else d2
        assert (d1 is None) == (d2 is None)
        if d1 is None:
            return x
        d1, d2 = np.squeeze(d1), np.squeeze(d2)
        assert d1.shape == d2.shape == (2,)
        assert d1[0] < d1[1] and d2[0] < d2[1]
        x = np.asarray(x, dtype=np.float64)
        shape = x.shape
        x = np.squeeze(x)
        ok = ((x >= d1[0]) | np.isclose(x, d1[0])) & ((x <= d1[1]) | np.isclose(x, d1[1]))
        assert np.all(ok), f"values outside the source domain: {x[~ok]}"
        return _affine(x, d1, d2).reshape(shape)

    def scale_back(self, x):
        return self.scale(x, d1=self._domain, d2=self._ref)

    def get_random(self, n: int = 1) -> np.ndarray:
        return self.scale(np.random.uniform(-1, 1, n))

    def _table_key(self):
        if self._domain is None:
            return ("leja", None)
        return ("leja", float(self._domain[0]), float(self._domain[1]))

    def __repr__(self) -> str:
        return f"Leja (domain = {self._domain})"


class Leja(Generator):
    """leja.py:207-343: one Leja1D per dim; default domains share a single instance."""

    def __init__(self, *, domains=None, dim: int = None):
        self._domains = None
        if domains is not None:
            Generator.__init__(self, [Leja1D(dm) for dm in domains])
            self._domains = np.asarray(domains, dtype=np.float64)
        elif dim is not None:
            Generator.__init__(self, [Leja1D()] * dim)
        else:
            raise ValueError("Must specify one of 'domains' or 'dim'.")

    def _map(self, x, src, dst):
        x = np.asarray(x, dtype=np.float64)
        shape = x.shape
        src, dst = np.asarray(src, dtype=np.float64), np.asarray(dst, dtype=np.float64)
        if src.ndim == 1:
            src, dst = src.reshape(1, 2), dst.reshape(1, 2)
        d = src.shape[0]
        if x.ndim == 0:
            x = x.reshape(1, 1)
        elif x.ndim == 1:
            x = x.reshape(1, -1) if x.shape[0] == d else x.reshape(-1, 1)
        ok = ((x >= src[:, 0]) | np.isclose(x, src[:, 0])) & ((x <= src[:, 1]) | np.isclose(x, src[:, 1]))
        assert np.all(ok), f"values outside the source domain: {x[~ok]}"
        x = (x - src[:, 0]) / (src[:, 1] - src[:, 0])
        return (x * (dst[:, 1] - dst[:, 0]) + dst[:, 0]).reshape(shape)

    def scale(self, x, d1=None, d2=None):
        if self._domains is None and d1 is None and d2 is None:
            return x
        d1 = np.array([[-1.0, 1.0]] * self.dim) if d1 is None else d1
        d2 = self._domains if d2 is None else d2
        return self._map(x, d1, d2)

    def scale_back(self, x):
        if self._domains is None:
            return x
        return self._map(x, self._domains, np.array([[-1.0, 1.0]] * self.dim))

    def __repr__(self) -> str:
        if self._domains is not None:
            return f"Leja (d = {self.dim}, domains = {self._domains.tolist()})"
        return f"Leja (d = {self.dim})"

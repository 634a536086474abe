"""Oracle restatement of the reference node generators and barycentric weights.

Follows src/smolyax/nodes/leja.py, src/smolyax/nodes/gausshermite.py, src/smolyax/nodes/base.py
and src/smolyax/barycentric.py:13-31.  Test infrastructure only.

Pinned by (tests/test_oracle_nodes.py):
  * Leja closed form: the sequence is 0, 1, -1 and then +-cos of bisected angles;
    every node is the cosine of a dyadic multiple of pi,
  * Leja quadrature weights integrate x^i exactly against dx/2 on [-1,1],
  * Gauss-Hermite nodes / weights from numpy.polynomial.hermite.hermgauss (the routine the
    reference calls, gausshermite.py:60,85); weights sum to 1 and reproduce N(0,1/2) moments,
  * barycentric weights against the textbook product formula; the normalised barycentric basis
    equals the Lagrange basis (checked against numpy.polynomial Lagrange interpolation).
"""

import numpy as np


class Leja1D:
    """leja.py:10-204. Nested nodes on [-1,1], optional affine map to `domain`."""

    is_nested = True
    _nodes = np.array([0, 1, -1, 1 / np.sqrt(2), -1 / np.sqrt(2)])

    def __init__(self, domain=None):
        self.domain = None if domain is None else np.asarray(domain, dtype=float)

    @classmethod
    def _ensure(cls, n):
        k = cls._nodes.shape[0]
        if n >= k:
            cls._nodes = np.append(cls._nodes, np.empty((n + 1 - k,)))
            for j in range(k, n + 1):
                if j % 2 == 0:
                    cls._nodes[j] = -cls._nodes[j - 1]
                else:
                    cls._nodes[j] = np.sqrt((cls._nodes[int((j + 1) / 2)] + 1) / 2)

    def scale(self, x):
        """leja.py:95-151 (reference domain [-1,1] -> self.domain)."""
        if self.domain is None:
            return x
        x = np.asarray(x, dtype=float)
        y = (x - (-1)) / (1 - (-1))
        return y * (self.domain[1] - self.domain[0]) + self.domain[0]

    def scale_back(self, x):
        if self.domain is None:
            return x
        x = np.asarray(x, dtype=float)
        y = (x - self.domain[0]) / (self.domain[1] - self.domain[0])
        return y * 2.0 + (-1.0)

    def __call__(self, n):
        self._ensure(n)
        return self.scale(self._nodes[: n + 1].copy())

    def get_quadrature_weights(self, n):
        """leja.py:83-93: Vandermonde solve against the moments of dx/2 on [-1,1]."""
        self._ensure(n)
        q = self._nodes[: n + 1]
        vm = np.vstack([q**i for i in range(n + 1)])
        rhs = np.array([(1 + (-1) ** i) / (2.0 * (i + 1)) for i in range(n + 1)])
        return np.linalg.solve(vm, rhs)


class GaussHermite1D:
    """gausshermite.py:10-164."""

    is_nested = False

    def __init__(self, mean=0.0, scaling=1.0):
        self.mean = mean
        self.scaling = scaling

    def scale(self, x):
        return self.mean + self.scaling * x

    def scale_back(self, x):
        return (x - self.mean) / self.scaling

    def __call__(self, n):
        return self.scale(np.polynomial.hermite.hermgauss(n + 1)[0])

    def get_quadrature_weights(self, n):
        return np.polynomial.hermite.hermgauss(n + 1)[1] / np.sqrt(np.pi)


class Generator:
    """base.py:131-262: a list of 1D generators, one per input dimension."""

    def __init__(self, gens):
        assert all(g.is_nested == gens[0].is_nested for g in gens)
        self.gens = list(gens)
        self.is_nested = gens[0].is_nested
        self.dim = len(gens)

    def __getitem__(self, i):
        return self.gens[i]

    def __iter__(self):
        return iter(self.gens)


def leja(dim=None, domains=None):
    """leja.py:212-244: shared 1D instance for default domains, one per dim otherwise."""
    if domains is not None:
        return Generator([Leja1D(dm) for dm in domains])
    g = Leja1D()
    return Generator([g] * dim)


def gauss_hermite(dim=None, mean=None, scaling=None):
    """gausshermite.py:173-208."""
    if dim is None:
        dim = len(scaling) if scaling is not None else len(mean)
    if mean is None and scaling is None:
        g = GaussHermite1D()
        return Generator([g] * dim)
    mean = np.zeros(dim) if mean is None else np.asarray(mean, dtype=float)
    scaling = np.ones(dim) if scaling is None else np.asarray(scaling, dtype=float)
    return Generator([GaussHermite1D(m, a) for m, a in zip(mean, scaling)])


def compute_weights(nodes):
    """barycentric.py:13-31: w_b = prod_{a != b} 1/(xi_a - xi_b) (reference sign convention)."""
    nodes = np.asarray(nodes, dtype=float)
    diffs = nodes[:, None] - nodes
    diffs = np.where(diffs == 0, 1, diffs)
    return np.prod(1 / diffs, axis=0)

"""CPU: pins of the oracle's extended-precision mode (oracle/smolyak.py, dtype=np.longdouble).

The extended mode is the reference of the GPU accuracy tests (tests/test_gpu_exact.py), so it
is checked here against exact rational arithmetic: the combination technique of the reference
(interpolation.py:264-390) written directly with fractions.Fraction over the same FP64 data (node
values, f at the nodes, quadrature weights), with the Lagrange basis
l_i(x) = prod_{a != i} (x - xi_a) / (xi_i - xi_a) and its derivative
l_i'(x) = l_i(x) sum_{a != i} 1 / (x - xi_a).  A planted mistake (FP64 barycentric weights in the
extended evaluation) must fail the same bound.
"""

import itertools as it
from fractions import Fraction

import numpy as np
import pytest

from oracle import indices as OI
from oracle import nodes as ON
from oracle import smolyak as OS

EXT = np.longdouble


def lagrange(xi, i, x):
    v = Fraction(1)
    for a, xa in enumerate(xi):
        if a != i:
            v *= (x - xa) / (xi[i] - xa)
    return v


def lagrange_d(xi, i, x):
    return lagrange(xi, i, x) * sum(Fraction(1) / (x - xa) for a, xa in enumerate(xi) if a != i)


class Exact:
    """The Smolyak interpolant of f's FP64 values, in exact rational arithmetic."""

    def __init__(self, k, t, gen, f, d_out):
        self.d = len(k)
        self.gen, self.f, self.d_out = gen, f, d_out
        n2nus, n2z = OI.non_zero_indices_and_zetas(k, t)
        self.terms = [(nu, z) for n in n2nus for nu, z in zip(n2nus[n], n2z[n])]
        self.zero = np.array([g(0)[0] for g in gen])

    def _blocks(self):
        for nu, z in self.terms:
            dims = [j for j, _ in nu]
            xis = [[Fraction(float(v)) for v in self.gen[j](deg)] for j, deg in nu]
            for mu in it.product(*[range(deg + 1) for _, deg in nu]):
                X = self.zero.copy()
                for q, j in enumerate(dims):
                    X[j] = self.gen[j](nu[q][1])[mu[q]]
                fv = [Fraction(float(v)) for v in np.asarray(self.f(X), dtype=np.float64).reshape(-1)]
                yield z, dims, xis, mu, fv, nu

    def __call__(self, x):
        out = []
        for p in x:
            y = [Fraction(0)] * self.d_out
            for z, dims, xis, mu, fv, _ in self._blocks():
                w = Fraction(z)
                for q, j in enumerate(dims):
                    w *= lagrange(xis[q], mu[q], Fraction(float(p[j])))
                y = [a + w * b for a, b in zip(y, fv)]
            out.append(y)
        return out

    def gradient(self, x):
        out = []
        for p in x:
            J = [[Fraction(0)] * self.d for _ in range(self.d_out)]
            for z, dims, xis, mu, fv, _ in self._blocks():
                xs = [Fraction(float(p[j])) for j in dims]
                for a, ja in enumerate(dims):
                    w = Fraction(z) * lagrange_d(xis[a], mu[a], xs[a])
                    for q in range(len(dims)):
                        if q != a:
                            w *= lagrange(xis[q], mu[q], xs[q])
                    for o in range(self.d_out):
                        J[o][ja] += w * fv[o]
            out.append(J)
        return out

    def integral(self):
        Q = [Fraction(0)] * self.d_out
        for z, dims, xis, mu, fv, nu in self._blocks():
            w = Fraction(z)
            for q, j in enumerate(dims):
                w *= Fraction(float(self.gen[j].get_quadrature_weights(nu[q][1])[mu[q]]))
            Q = [a + w * b for a, b in zip(Q, fv)]
        return Q


def as_float(v):
    return np.array(v, dtype=object).astype(np.float64)


def ext_err(a_ext, exact):
    """max |a - exact| / max |exact|, the difference taken in the extended type from the
    float64 pair of each exact value (hi + lo represents it to ~1e-32)."""
    e = np.array(exact, dtype=object)
    hi = e.astype(np.float64)
    lo = (e - np.vectorize(Fraction)(hi)).astype(np.float64)
    diff = (np.asarray(a_ext, dtype=EXT) - hi.astype(EXT)) - lo.astype(EXT)
    return float(np.max(np.abs(diff)) / np.max(np.abs(hi)))


CASES = [
    # d, d_out, target node count, node family
    (2, 2, 14, "leja"),
    (3, 1, 25, "leja"),
    (2, 2, 20, "gh"),
]


def build(d, d_out, m, kind):
    k = np.log((2 + np.arange(d)) / np.log(2))
    gen = ON.leja(dim=d) if kind == "leja" else ON.gauss_hermite(dim=d)
    t = OI.find_approximate_threshold(k, m, gen.is_nested)
    A = np.random.default_rng(d + d_out).standard_normal((d_out, d))
    f = lambda x: np.exp(np.sin(A @ np.asarray(x)))
    return OS.Interpolant(d_out=d_out, node_gen=gen, k=k, t=t, f=f), Exact(k, t, gen, f, d_out)


@pytest.mark.parametrize("d,d_out,m,kind", CASES)
def test_extended_mode_is_exact(d, d_out, m, kind):
    ref, ex = build(d, d_out, m, kind)
    x = np.random.default_rng(7).uniform(-0.9, 0.9, (3, d))
    y_exact = ex(x)
    assert ext_err(ref(x, dtype=EXT), y_exact) < 1e-17
    # FP64 mode: the reference's algorithm, within its own rounding (far above the extended mode)
    assert ext_err(ref(x), y_exact) < 1e-13
    assert ext_err(ref.gradient(x, dtype=EXT), ex.gradient(x)) < 1e-17
    assert ext_err(ref.integral(dtype=EXT), ex.integral()) < 1e-17


def test_extended_mode_planted_fp64_weights_fail():
    """Extended evaluation with the FP64 barycentric weights (the mistake of not recomputing them)
    misses the exact values by the weights' FP64 rounding: the 1e-17 bound catches it."""
    ref, ex = build(3, 1, 25, "leja")
    x = np.random.default_rng(7).uniform(-0.9, 0.9, (3, 3))
    y_exact = ex(x)
    for g in ref.groups.values():
        g["weights_ext"] = g["weights"].astype(EXT)
    assert ext_err(ref(x, dtype=EXT), y_exact) > 1e-17

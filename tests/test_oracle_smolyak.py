"""CPU: pins of the oracle's Smolyak interpolant (oracle/smolyak.py) and of the CPU baseline port.

The Smolyak operator reproduces every polynomial in P_Lambda = span{x^nu : nu in Lambda} (the
reference's own check, tests/test_smolyak.py:7-26), so for p = sum_nu c_nu prod_j P_{nu_j}(x_j)
with tensor Legendre (Leja nodes on [-1,1]) or Hermite (Gauss-Hermite nodes) polynomials:
  * I[p](x) = p(x) at random points (value path, a3/a4/a7),
  * dI[p]/dx = dp/dx from the analytic derivatives (gradient path, a8),
  * Q[p] = c_0, the orthogonality integral (quadrature path, a9),
plus exact interpolation at the sparse-grid nodes, NaN gradients at grid coordinates, constant
reproduction (sum of zeta = 1) and the torch-CPU baseline port against the oracle.
A dropped term, a wrong sign of a zeta, a transposed block axis or a wrong derivative factor makes
the reproduced polynomial differ by O(1).
"""

import math

import numpy as np
import pytest

from oracle import indices as OI
from oracle import nodes as ON
from oracle import smolyak as OS


def k_rule(d):
    return np.log((2 + np.arange(d)) / np.log(2))


def poly_family(kind, scaling=1.0):
    if kind == "leja":
        P = np.polynomial.legendre.Legendre
        return (lambda v: P.basis(v)), (lambda v: P.basis(v).deriv()), 1.0
    H = np.polynomial.hermite.Hermite
    return (lambda v: H.basis(v)), (lambda v: H.basis(v).deriv()), scaling


class TensorPoly:
    """p_m(x) = sum_i c[i, m] prod_{(j, v) in nus[i]} B_v(x_j / s)."""

    def __init__(self, kind, nus, d_in, d_out, seed, scaling=1.0):
        self.basis, self.dbasis, self.s = poly_family(kind, scaling)
        self.nus = list(nus)
        self.d_in, self.d_out = d_in, d_out
        rng = np.random.default_rng(seed)
        self.c = rng.standard_normal((len(self.nus), d_out))
        self.c0 = self.c[self.nus.index(())]

    def _factors(self, x):
        x = np.atleast_2d(x) / self.s
        cache = {}

        def fac(j, v, der=False):
            key = (j, v, der)
            if key not in cache:
                cache[key] = (self.dbasis(v) if der else self.basis(v))(x[:, j])
            return cache[key]
        return x.shape[0], fac

    def __call__(self, x):
        P, fac = self._factors(x)
        y = np.zeros((P, self.d_out))
        for i, nu in enumerate(self.nus):
            w = np.ones(P)
            for j, v in nu:
                w = w * fac(j, v)
            y += w[:, None] * self.c[i]
        return y

    def point(self, x):
        return self(np.asarray(x)[None, :])[0]

    def grad(self, x):
        P, fac = self._factors(x)
        J = np.zeros((P, self.d_out, self.d_in))
        for i, nu in enumerate(self.nus):
            for a, (ja, va) in enumerate(nu):
                w = fac(ja, va, True) / self.s
                for b, (jb, vb) in enumerate(nu):
                    if b != a:
                        w = w * fac(jb, vb)
                J[:, :, ja] += w[:, None] * self.c[i]
        return J


CASES = [("leja", 3, 6.0), ("leja", 5, 5.5), ("gh", 3, 5.0), ("gh", 4, 4.8)]


def build(kind, d, t, d_out=2, seed=0):
    k = k_rule(d)
    nus = OI.indexset(k, t)
    if kind == "leja":
        gen, sc = ON.leja(dim=d), 1.0
    else:
        sc = math.sqrt(2.0)
        gen = ON.gauss_hermite(dim=d, scaling=np.full(d, sc))
    p = TensorPoly(kind, nus, d, d_out, seed, scaling=sc)
    ip = OS.Interpolant(d_out=d_out, node_gen=gen, k=k, t=t, f=p.point)
    return ip, p, nus


def sample(kind, d, n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, (n, d)) if kind == "leja" else rng.standard_normal((n, d))


@pytest.mark.parametrize("kind,d,t", CASES)
def test_polynomial_reproduction(kind, d, t):
    ip, p, nus = build(kind, d, t)
    assert len(nus) > 10
    x = sample(kind, d, 64, 1)
    y, ref = ip(x), p(x)
    assert np.max(np.abs(y - ref)) <= 1e-11 * np.max(np.abs(ref))


@pytest.mark.parametrize("kind,d,t", CASES)
def test_gradient_of_reproduced_polynomial(kind, d, t):
    ip, p, _ = build(kind, d, t)
    x = sample(kind, d, 32, 2)
    J, ref = ip.gradient(x), p.grad(x)
    assert J.shape == (32, p.d_out, d)
    assert np.max(np.abs(J - ref)) <= 1e-10 * np.max(np.abs(ref))


@pytest.mark.parametrize("kind,d,t", CASES)
def test_integral_is_constant_coefficient(kind, d, t):
    ip, p, _ = build(kind, d, t)
    Q = ip.integral()
    assert np.max(np.abs(Q - p.c0)) <= 1e-12 * np.max(np.abs(p.c))


def test_exact_at_grid_nodes_and_nan_gradient():
    d, t = 4, 5.5
    k = k_rule(d)
    gen = ON.leja(dim=d)
    f = lambda x: np.array([np.exp(-np.sum(x / (1 + np.arange(d)) ** 2)), np.sin(x[0] + 2 * x[1])])
    ip = OS.Interpolant(d_out=2, node_gen=gen, k=k, t=t, f=f)
    nus = OI.indexset(k, t)
    X = np.zeros((len(nus), d))
    for i, nu in enumerate(nus):
        for j, v in nu:
            X[i, j] = gen[j](v)[v]
    F = np.stack([f(x) for x in X])
    assert np.max(np.abs(ip(X) - F)) <= 4e-15 * np.max(np.abs(F))
    x = np.array([[0.3, -0.2, 0.1, 0.05]])
    x[0, 1] = gen[1](2)[2]  # a grid coordinate in dimension 1
    J = ip.gradient(x)
    assert np.isnan(J[0, :, 1]).all()


def test_constant_reproduction_and_zeta_sum():
    d, t = 6, 6.0
    k = k_rule(d)
    _, n2z = OI.non_zero_indices_and_zetas(k, t)
    assert sum(sum(z) for z in n2z.values()) == 1
    ip = OS.Interpolant(d_out=1, node_gen=ON.leja(dim=d), k=k, t=t, f=lambda x: 2.5)
    y = ip(np.random.default_rng(0).uniform(-1, 1, (10, d)))
    assert np.max(np.abs(y - 2.5)) < 1e-13
    assert abs(ip.integral()[0] - 2.5) < 1e-13


@pytest.mark.parametrize("kind", ["leja", "gh"])
def test_cpu_baseline_port_matches_oracle(kind):
    torch = pytest.importorskip("torch")
    from oracle import cpu_baseline as CB

    ip, p, _ = build(kind, 4, 5.0, d_out=3)
    x = sample(kind, 4, 50, 3)
    y = CB.TorchCPUInterpolant(ip)(x)
    y = y.numpy() if isinstance(y, torch.Tensor) else np.asarray(y)
    assert np.max(np.abs(y - ip(x))) <= 1e-13 * np.max(np.abs(ip(x)))

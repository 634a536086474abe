"""CPU: pins of the oracle's node generators and barycentric weights (oracle/nodes.py).

Each check takes an independent route to the same numbers, chosen so that a wrong index, sign or
normalisation in the oracle fails it:
  * Leja nodes against the cosines of explicitly bisected angles (leja.py:17,48-57 builds them by
    the half-angle recurrence on the node values),
  * Leja quadrature weights against exact integration of the Lagrange basis polynomials
    (numpy.polynomial), and exactness on Legendre polynomials,
  * Gauss-Hermite nodes and weights against the Golub-Welsch eigen-decomposition of the Hermite
    Jacobi matrix (the reference calls numpy's hermgauss, gausshermite.py:60,85), and the moments
    of the normalised measure,
  * barycentric weights against the product formula (barycentric.py:29-31) and the normalised
    barycentric basis against the Lagrange basis fitted by numpy.polynomial.
"""

import math

import numpy as np
import pytest

from oracle import nodes as ON


def leja_by_angles(n):
    """theta = pi/2, 0, pi, pi/4, 3pi/4 for the first five nodes; then odd j:
    theta_j = theta_{(j+1)/2} / 2 and even j: theta_j = pi - theta_{j-1}; node = cos(theta)."""
    th = [math.pi / 2, 0.0, math.pi, math.pi / 4, 3 * math.pi / 4]
    for j in range(5, n + 1):
        th.append(th[(j + 1) // 2] / 2 if j % 2 == 1 else math.pi - th[j - 1])
    return np.cos(np.array(th[: n + 1]))


def test_leja_nodes_closed_form():
    n = 40
    xi = ON.Leja1D()(n)
    ref = leja_by_angles(n)
    assert xi.shape == (n + 1,)
    assert np.max(np.abs(xi - ref)) < 4e-16
    assert xi[0] == 0.0 and xi[1] == 1.0 and xi[2] == -1.0
    assert len(np.unique(xi)) == n + 1
    # first nodes written out
    c8, c38 = math.cos(math.pi / 8), math.cos(3 * math.pi / 8)
    np.testing.assert_allclose(xi[3:9], [math.sqrt(0.5), -math.sqrt(0.5), c8, -c8, c38, -c38],
                               rtol=0, atol=2e-16)


def test_leja_nested_prefixes():
    g = ON.Leja1D()
    full = g(20)
    for n in range(21):
        assert np.array_equal(g(n), full[: n + 1])


def test_leja_domain_map():
    g = ON.Leja1D(domain=[2.0, 5.0])
    x = g(6)
    np.testing.assert_allclose(x, 2.0 + (leja_by_angles(6) + 1) * 1.5, rtol=0, atol=1e-15)
    np.testing.assert_allclose(g.scale_back(x), leja_by_angles(6), atol=1e-15)


@pytest.mark.parametrize("n", [0, 1, 2, 5, 9])
def test_leja_quadrature_weights(n):
    xi = ON.Leja1D()(n)
    q = ON.Leja1D().get_quadrature_weights(n)
    # independent: q_i = int L_i dx/2 with L_i fitted through the one-hot data
    for i in range(n + 1):
        e = np.zeros(n + 1)
        e[i] = 1.0
        L = np.polynomial.Polynomial.fit(xi, e, n, domain=[-1, 1], window=[-1, 1]) if n else \
            np.polynomial.Polynomial([1.0])
        Li = L.integ()
        assert abs((Li(1.0) - Li(-1.0)) / 2 - q[i]) < 1e-12
    # exact on Legendre polynomials: int P_k dx/2 = delta_k0
    for k in range(n + 1):
        Pk = np.polynomial.legendre.Legendre.basis(k)
        assert abs(np.dot(q, Pk(xi)) - (1.0 if k == 0 else 0.0)) < 1e-12


def golub_welsch_hermite(m):
    """Nodes/weights of int f(t) exp(-t^2) dt / sqrt(pi) from the Jacobi matrix of the
    physicists' Hermite recurrence (off-diagonal sqrt(k/2))."""
    off = np.sqrt(np.arange(1, m) / 2.0)
    J = np.diag(off, 1) + np.diag(off, -1)
    lam, V = np.linalg.eigh(J)
    return lam, V[0, :] ** 2


@pytest.mark.parametrize("n", [0, 1, 4, 9, 15])
def test_gauss_hermite_against_golub_welsch(n):
    g = ON.GaussHermite1D()
    t, w = golub_welsch_hermite(n + 1)
    np.testing.assert_allclose(g(n), t, rtol=0, atol=1e-13)
    np.testing.assert_allclose(g.get_quadrature_weights(n), w, rtol=0, atol=1e-14)


def test_gauss_hermite_moments_and_scaling():
    n = 7
    g = ON.GaussHermite1D(mean=0.5, scaling=math.sqrt(2.0))
    x, q = g(n), g.get_quadrature_weights(n)
    assert abs(q.sum() - 1.0) < 1e-14
    # with scaling sqrt(2) the measure is N(mean, 1): E[x - m] = 0, E[(x-m)^2] = 1, E[(x-m)^4] = 3
    c = x - 0.5
    assert abs(np.dot(q, c)) < 1e-14
    assert abs(np.dot(q, c ** 2) - 1.0) < 1e-13
    assert abs(np.dot(q, c ** 4) - 3.0) < 1e-12
    np.testing.assert_allclose(g.scale_back(x), ON.GaussHermite1D()(n), atol=1e-15)


@pytest.mark.parametrize("kind,n", [("leja", 4), ("leja", 9), ("gh", 3), ("gh", 8)])
def test_barycentric_weights_and_lagrange_basis(kind, n):
    xi = ON.Leja1D()(n) if kind == "leja" else ON.GaussHermite1D()(n)
    w = ON.compute_weights(xi)
    for b in range(n + 1):  # product formula, the reference's sign: prod_a 1 / (xi_a - xi_b)
        ref = 1.0
        for a in range(n + 1):
            if a != b:
                ref /= xi[a] - xi[b]
        assert abs(w[b] - ref) <= 1e-13 * abs(ref)
    rng = np.random.default_rng(n)
    x = rng.uniform(xi.min(), xi.max(), 25)
    bas = (w / (x[:, None] - xi)) / np.sum(w / (x[:, None] - xi), axis=1, keepdims=True)
    for mu in range(n + 1):
        e = np.zeros(n + 1)
        e[mu] = 1.0
        L = np.polynomial.Polynomial.fit(xi, e, n) if n else np.polynomial.Polynomial([1.0])
        np.testing.assert_allclose(bas[:, mu], L(x), rtol=0, atol=1e-9)

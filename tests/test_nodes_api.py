"""CPU: the node generators' sampling and domain maps on smolyax_amd.nodes, restating the
reference's tests/test_nodes.py:5-24 (get_random shapes, scale(scale_back(x)) == x) for default
generators (d = 1..3) and for custom domains, means and scalings (d = 2..10), and checking that
samples come from the measure the quadrature integrates against (base.py:95-110): uniform on each
Leja domain, N(mean, scaling^2 / 2) for Gauss-Hermite (gausshermite.py:145)."""

import numpy as np
import pytest

from smolyax_amd import nodes as SN


def default_generators():
    out = []
    for d in range(1, 4):
        out += [SN.Leja(dim=d), SN.GaussHermite(dim=d)]
    return out


def custom_generators(rng, n=5, dmin=2, dmax=10):
    out = []
    for _ in range(n):
        d = int(rng.integers(dmin, dmax + 1))
        out.append(SN.Leja(domains=np.sort(rng.random((d, 2)), axis=1)))
        out.append(SN.GaussHermite(rng.standard_normal(d), rng.random(d)))
    return out


def test_nodes_scaling_default():
    rng = np.random.default_rng(0)
    for gen in default_generators():
        n = int(rng.integers(1, 5))
        x = gen.get_random(n)
        assert x.shape == (n, gen.dim)
        assert np.allclose(x, gen.scale(gen.scale_back(x)))


def test_nodes_scaling_custom():
    rng = np.random.default_rng(1)
    for gen in custom_generators(rng):
        for n in range(0, 5):
            x = gen.get_random(n)
            assert x.shape == ((gen.dim,) if n == 0 else (n, gen.dim))
            assert np.allclose(x, gen.scale(gen.scale_back(x)))


@pytest.mark.parametrize("kind", ["leja", "gh"])
def test_get_random_follows_the_measure(kind):
    rng = np.random.default_rng(2)
    d = 4
    if kind == "leja":
        dom = np.sort(rng.random((d, 2)), axis=1) + [0.0, 0.5]
        gen = SN.Leja(domains=dom)
    else:
        mean, sc = rng.standard_normal(d), rng.random(d) + 0.5
        gen = SN.GaussHermite(mean, sc)
    np.random.seed(3)  # the generators draw from numpy's global stream, as the reference does
    x = gen.get_random(20000)
    assert x.shape == (20000, d)
    if kind == "leja":
        assert np.all(x >= dom[:, 0]) and np.all(x <= dom[:, 1])
        assert np.allclose(x.mean(0), dom.mean(1), atol=0.02)
        assert np.allclose(x.var(0), (dom[:, 1] - dom[:, 0]) ** 2 / 12, rtol=0.05)
    else:
        assert np.allclose(x.mean(0), mean, atol=0.03)
        assert np.allclose(x.var(0), sc**2 / 2, rtol=0.05)


def test_scale_maps_reference_nodes_to_the_domain():
    dom = np.array([[2.0, 5.0], [-3.0, -1.0]])
    gen = SN.Leja(domains=dom)
    ref = SN.Leja(dim=2)
    for j in range(2):
        xi = gen[j](6)
        assert np.allclose(gen[j].scale_back(xi), ref[j](6))
        assert np.all(xi >= dom[j, 0]) and np.all(xi <= dom[j, 1])

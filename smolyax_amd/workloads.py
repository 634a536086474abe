"""Synthetic workloads of the benchmark configurations (seeded), shared by tests and bench.py.

k_j = log((2 + j) / log 2) (the reference README's rule); t from find_approximate_threshold so
that the node count hits the target; points uniform on [-1, 1]^d (Leja) or N(0, I) (GH) from a
seeded default_rng, following benchmarking/benchmark.py:160,187 of the reference.
"""

import numpy as np


def k_rule(d):
    return np.array([np.log((2 + j) / np.log(2)) for j in range(d)])


def f_rational(d):
    """1 / (1 + sum_j c_j x_j), c_j = 0.5 / (j+1)^2 (scalar)."""
    c = 0.5 / (np.arange(d) + 1.0) ** 2

    def f(x):
        x = np.asarray(x)
        return 1.0 / (1.0 + x @ c)
    return f


def f_exp(d):
    """exp(-sum_j c_j x_j), c_j = 1 / (j+1)^2 (scalar)."""
    c = 1.0 / (np.arange(d) + 1.0) ** 2

    def f(x):
        x = np.asarray(x)
        return np.exp(-(x @ c))
    return f


def f_sin(d, d_out, seed=1):
    """f_m(x) = sin(a_m . x + b_m), a_mj ~ N(0,1)/(j+1)^2, b_m ~ U(0, 2 pi)."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((d_out, d)) / (np.arange(d) + 1.0) ** 2
    b = rng.uniform(0, 2 * np.pi, d_out)

    def f(x):
        x = np.asarray(x)
        return np.sin(x @ A.T + b)
    return f


def f_tanh(d, d_out, seed=1):
    """f_m(x) = tanh(a_m . x), a_mj ~ N(0,1)/(j+1)^1.5."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((d_out, d)) / (np.arange(d) + 1.0) ** 1.5

    def f(x):
        x = np.asarray(x)
        return np.tanh(x @ A.T)
    return f


CONFIGS = {
    # name: (d_in, d_out, target nodes, points, f factory)
    "c1_rational_d10": dict(d_in=10, d_out=1, m=1000, P=1024, f=lambda: f_rational(10)),
    "c2_exp_d1000": dict(d_in=1000, d_out=1, m=10000, P=16384, f=lambda: f_exp(1000)),
    "c3_sin_d100_o4096": dict(d_in=100, d_out=4096, m=10000, P=4096, f=lambda: f_sin(100, 4096)),
    "c4_tanh_d200_o256": dict(d_in=200, d_out=256, m=5000, P=8192, f=lambda: f_tanh(200, 256),
                              vjp=True, oracle_points=8),
}


def points_uniform(P, d, seed=0):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, (P, d))


def serving_specs(n=32, seed=0):
    """Resident interpolants of the serving workload: Gauss-Hermite nodes scaled so their measure
    is N(0, I), d_in log-uniform in [8, 500], node count log-uniform in [1e2, 2e4], d_out
    log-uniform in [1, 1024], f_m(x) = sin(a_m . x + b_m) with a_mj ~ N(0,1)/(j+1)^2."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n):
        d = int(round(np.exp(rng.uniform(np.log(8), np.log(500)))))
        m = int(round(np.exp(rng.uniform(np.log(1e2), np.log(2e4)))))
        d_out = int(round(np.exp(rng.uniform(np.log(1), np.log(1024)))))
        specs.append(dict(name=f"ip{i:02d}", d_in=d, nodes=m, d_out=d_out, fseed=1000 + i))
    return specs


def build_serving_interpolants(specs, device=None, progress=None):
    from . import indices, nodes
    from .interpolation import SmolyakBarycentricInterpolator

    out = {}
    for sp in specs:
        d = sp["d_in"]
        k = k_rule(d)
        ng = nodes.GaussHermite(scaling=np.full(d, np.sqrt(2.0)))
        t = indices.find_approximate_threshold(k, sp["nodes"], False, accuracy=0.05)
        ip = SmolyakBarycentricInterpolator(d_out=sp["d_out"], node_gen=ng, k=k, t=t,
                                            device=device, precision="mixed")
        f = f_sin(d, sp["d_out"], seed=sp["fseed"])
        ip.set_f(f=lambda X, f=f: f(X), vectorized=True)
        out[sp["name"]] = ip
        if progress is not None:
            progress(f"serve: built {sp['name']} (d_in={d}, k_raw={ip.layout.k_raw}, d_out={sp['d_out']})")
    return out

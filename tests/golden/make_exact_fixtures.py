"""Write the extended-precision fixtures exact_<workload>.npz used by tests/test_gpu_exact.py.

    python tests/golden/make_exact_fixtures.py [workload ...]

For each benchmark workload (d_in, k rule, node count and f of the configuration, points from
the workload's seed-0 point set) the oracle (oracle/smolyak.py, the restatement of the reference's
combination technique) evaluates __call__, gradient and integral twice: in FP64, as the reference
does, and in the oracle's extended mode (np.longdouble: barycentric weights recomputed from the
FP64 nodes, every product and sum in the 64-bit-significand type), which is the exact polynomial
interpolant of the FP64 data to ~1e-19 relative to the combination's partial sums.

Stored per workload (all float64 unless noted):
  x           (32, d_in)  the first 32 points of the workload's point set
  cols        (n_cols,)   int64 output columns covered (all of them unless d_out > 256)
  y_hi, y_lo  (32, n_cols) exact values as a float64 pair (y = hi + lo)
  y_f64       (32, n_cols) the FP64 oracle's values (the reference's algorithm in FP64)
  J_hi        (nJ, n_cols_J, d_in) exact gradient at the first nJ points, columns cols[:n_cols_J]
  J_f64       same shape, FP64 oracle
  q_hi, q_f64 (n_cols,)   integral
  w, g_hi, g_f64          (d_out,), (32, d_in): the adjoint g = sum_m w_m dy_m/dx of the
                          configuration with a weight vector (only where the workload has one)
  f_first     (16, n_cols) f at the first 16 sparse-grid nodes, to show that the machine that
                          rebuilds the interpolant evaluates f to the same values
"""

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import indices as OI  # noqa: E402
from oracle import nodes as ON  # noqa: E402
from oracle import smolyak as OS  # noqa: E402
from smolyax_amd import workloads as W  # noqa: E402

EXT = np.longdouble
N_POINTS = 32
SPEC = {
    # workload: (points of the gradient, output columns kept (None: all), columns of the gradient,
    #            adjoint weight vector)
    "c1_rational_d10": (8, None, None, False),
    "c2_exp_d1000": (8, None, None, False),
    "c3_sin_d100_o4096": (4, 64, 64, False),
    "c4_tanh_d200_o256": (8, None, 32, True),
}


def split(v):
    hi = np.asarray(v, dtype=np.float64)
    lo = np.asarray(np.asarray(v, dtype=EXT) - hi.astype(EXT), dtype=np.float64)
    return hi, lo


def columns(name, d_out, keep):
    if keep is None or keep >= d_out:
        return np.arange(d_out, dtype=np.int64)
    return np.sort(np.random.default_rng(5).choice(d_out, keep, replace=False)).astype(np.int64)


def make(name):
    cfg = W.CONFIGS[name]
    d, d_out = cfg["d_in"], cfg["d_out"]
    nJ, keep, keepJ, adjoint = SPEC[name]
    f = cfg["f"]()
    k = W.k_rule(d)
    t = OI.find_approximate_threshold(k, cfg["m"], True)
    cols = columns(name, d_out, keep)
    full = len(cols) == d_out
    fc = f if full else (lambda X: np.asarray(f(X))[..., cols])
    t0 = time.perf_counter()
    ref = OS.Interpolant(d_out=len(cols), node_gen=ON.leja(dim=d), k=k, t=t, f=fc)
    x = W.points_uniform(cfg["P"], d, seed=0)[:N_POINTS]
    out = {"x": x, "cols": cols, "t": np.array(t)}
    y_ext = ref(x, dtype=EXT)
    out["y_hi"], out["y_lo"] = split(y_ext)
    out["y_f64"] = ref(x)
    nc = len(cols) if keepJ is None else min(keepJ, len(cols))
    J_ext = ref.gradient(x[: N_POINTS if adjoint else nJ], dtype=EXT)
    out["J_hi"], _ = split(J_ext[:nJ, :nc])
    out["J_f64"] = ref.gradient(x[:nJ])[:, :nc]
    out["q_hi"], _ = split(ref.integral(dtype=EXT))
    out["q_f64"] = ref.integral()
    if adjoint:
        assert full
        w = np.random.default_rng(2).standard_normal(d_out)
        out["w"] = w
        out["g_hi"], _ = split(np.einsum("m,pmj->pj", w.astype(EXT), J_ext))
        out["g_f64"] = np.einsum("m,pmj->pj", w, ref.gradient(x))
    nus = OI.indexset(k, t)[:16]
    zero = np.zeros(d)
    gen = ON.leja(dim=d)
    X = np.tile(zero, (len(nus), 1))
    for r, nu in enumerate(nus):
        for j, kk in nu:
            X[r, j] = gen[j](kk)[kk]
    out["f_first"] = np.stack([np.asarray(fc(X[r]), dtype=np.float64).reshape(-1)
                               for r in range(len(nus))])
    e = np.max(np.abs(out["y_f64"] - y_ext)) / np.max(np.abs(y_ext))
    print(f"{name}: |Lambda| {ref.n_f_evals}, {len(cols)} outputs, FP64 oracle vs exact "
          f"{float(e):.2e} ({time.perf_counter() - t0:.1f} s)", flush=True)
    np.savez_compressed(os.path.join(HERE, f"exact_{name}.npz"), **out)


if __name__ == "__main__":
    for name in sys.argv[1:] or list(SPEC):
        make(name)

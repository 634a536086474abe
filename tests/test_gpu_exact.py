"""GPU: every evaluation path against the exact interpolant, on the benchmark workloads.

Fixtures tests/golden/exact_<workload>.npz (written by tests/golden/make_exact_fixtures.py) hold,
for 32 points of each workload with the workload's own f, the exact polynomial interpolant of the
FP64 data (the oracle's extended-precision evaluation of the reference's combination technique,
itself pinned against exact rational arithmetic in tests/test_oracle_extended.py) and the FP64
oracle's values (the reference's algorithm in FP64).  The interpolant is rebuilt here at the
workload's full size (d_in, node count, d_out) and evaluated through the library.

Errors are max |result - exact| / max |exact| over the points and the covered outputs.  Bounds:

- newton-fp64 (the Newton form in FP64; no cancellation between terms): NEWTON_FP64_TOL;
- fp64 (the reference's combination technique in FP64): within FP64_FACTOR times the FP64
  oracle's own distance from the exact values.  The combination technique adds block values
  whose partial sums reach sum |zeta_nu| |y| (1e4 to 2.4e4 here), so any FP64 evaluation of it
  rounds at 1e-12 to 5e-11 of max |y| at these sizes, the reference's included;
- mixed (FP32 products, FP64 node differences and accumulation): MIXED_TOL = 1e-6;
- bf16 (BF16 operands on the matrix cores, FP32 partial sums flushed into FP64): BF16_TOL;
- gradient (FP64) to GRAD_TOL; the integral (the same zeta-weighted combination of block sums
  as the FP64 barycentric path, so the same rounding) within FP64_FACTOR times the FP64 oracle's
  own distance, at least 1e-14 of max |y|; the adjoint of the workload with a weight vector, which
  is also checked against central finite differences of w . y with h = 1e-6 in FP64 (relative
  1e-6: the differences' own rounding, ~1e-16 |w . y| / h, and truncation, ~h^2).
"""

import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
WORKLOADS = ["c1_rational_d10", "c2_exp_d1000", "c3_sin_d100_o4096", "c4_tanh_d200_o256"]

NEWTON_FP64_TOL = 1e-13
FP64_FACTOR = 4.0
MIXED_TOL = 1e-6
BF16_TOL = 3e-3
GRAD_TOL = 1e-12


def err(a, hi, lo=None):
    a = np.asarray(a, dtype=np.float64)
    d = (a - hi) - (lo if lo is not None else 0.0)
    return float(np.max(np.abs(d)) / np.max(np.abs(hi)))


@pytest.fixture(scope="module", params=WORKLOADS)
def case(request, gpu):
    from smolyax_amd import indices as SI
    from smolyax_amd import nodes as SN
    from smolyax_amd import workloads as W
    from smolyax_amd.interpolation import SmolyakBarycentricInterpolator

    name = request.param
    fx = dict(np.load(os.path.join(GOLDEN, f"exact_{name}.npz")))
    cfg = W.CONFIGS[name]
    d, d_out = cfg["d_in"], cfg["d_out"]
    f = cfg["f"]()
    k = W.k_rule(d)
    t = SI.find_approximate_threshold(k, cfg["m"], True)
    assert t == float(fx["t"])
    print(f"exact-value workload {name}: building d_in={d} d_out={d_out}", flush=True)
    ip = SmolyakBarycentricInterpolator(d_out=d_out, node_gen=SN.Leja(dim=d), k=k, t=t,
                                        memory_limit=16.0)
    fe = ip.set_f(f=lambda X: f(X), vectorized=True)
    # f at the first grid nodes on this machine against the values the fixture was made with
    from smolyax_amd import indices as _I

    nus = _I.indexset(k, t)[:16]
    here = np.stack([np.asarray(fe[nu], dtype=np.float64).reshape(-1)[fx["cols"]] for nu in nus])
    df = float(np.max(np.abs(here - fx["f_first"])) / np.max(np.abs(fx["f_first"])))
    print(f"{name}: f at the first grid nodes differs from the fixture's by {df:.1e}", flush=True)
    assert df <= 1e-15
    return dict(name=name, ip=ip, fx=fx, x=torch.as_tensor(fx["x"], device="cuda"))


@pytest.mark.parametrize("prec", ["newton-fp64", "fp64", "mixed", "bf16"])
def test_eval_against_exact(case, prec):
    ip, fx, name = case["ip"], case["fx"], case["name"]
    cols = fx["cols"]
    y = ip.evaluate(case["x"], prec).cpu().numpy()[:, cols]
    e = err(y, fx["y_hi"], fx["y_lo"])
    e64 = err(fx["y_f64"], fx["y_hi"], fx["y_lo"])
    print(f"{name} {prec}: {e:.2e} from the exact interpolant (FP64 oracle: {e64:.2e})",
          flush=True)
    if prec == "newton-fp64":
        assert e <= NEWTON_FP64_TOL
    elif prec == "fp64":
        assert e <= FP64_FACTOR * e64
    elif prec == "mixed":
        assert e <= MIXED_TOL
    else:
        assert e <= BF16_TOL


def test_gradient_against_exact(case):
    ip, fx, name = case["ip"], case["fx"], case["name"]
    nJ, nc = fx["J_hi"].shape[:2]
    J = ip.gradient(case["x"][:nJ]).cpu().numpy()[:, fx["cols"][:nc]]
    e = err(J, fx["J_hi"])
    e64 = err(fx["J_f64"], fx["J_hi"])
    print(f"{name} gradient: {e:.2e} from the exact Jacobian (FP64 oracle: {e64:.2e})", flush=True)
    assert not np.isnan(J).any()
    assert e <= GRAD_TOL


def test_integral_against_exact(case):
    ip, fx, name = case["ip"], case["fx"], case["name"]
    Q = ip.integral().cpu().numpy()[fx["cols"]]
    # relative to the size of the values: the integral of an odd f (the tanh workload) is close to zero
    scale = float(np.max(np.abs(fx["y_hi"])))
    e = float(np.max(np.abs(Q - fx["q_hi"]))) / scale
    e64 = float(np.max(np.abs(fx["q_f64"] - fx["q_hi"]))) / scale
    print(f"{name} integral: {e:.2e} of max |y| from the exact value (FP64 oracle: {e64:.2e})",
          flush=True)
    assert e <= max(FP64_FACTOR * e64, 1e-14)


@pytest.mark.parametrize("prec", ["newton-fp64", "mixed"])
def test_adjoint_against_exact(case, prec):
    ip, fx, name = case["ip"], case["fx"], case["name"]
    if "w" not in fx:
        pytest.skip("workload without an adjoint weight vector")
    x = case["x"]
    v = torch.as_tensor(np.tile(fx["w"], (x.shape[0], 1)), device="cuda")
    g = ip.vjp(x, v, precision=prec).cpu().numpy()
    e = err(g, fx["g_hi"])
    e64 = err(fx["g_f64"], fx["g_hi"])
    print(f"{name} adjoint {prec}: {e:.2e} from the exact value (FP64 oracle: {e64:.2e})",
          flush=True)
    assert e <= (GRAD_TOL if prec == "newton-fp64" else MIXED_TOL)


def test_adjoint_finite_differences(case):
    """Central differences of w . y (h = 1e-6, FP64 Newton form) at 8 of the points, every input
    dim, against the FP64 adjoint: the configuration's own finite-difference check at its size."""
    ip, fx, name = case["ip"], case["fx"], case["name"]
    if "w" not in fx:
        pytest.skip("workload without an adjoint weight vector")
    x = fx["x"][:8]
    P, d = x.shape
    h = 1e-6
    E = np.eye(d) * h
    Xp = (x[:, None, :] + E[None]).reshape(-1, d)
    Xm = (x[:, None, :] - E[None]).reshape(-1, d)
    w = fx["w"]
    yp = ip.evaluate(Xp, "newton-fp64").cpu().numpy() @ w
    ym = ip.evaluate(Xm, "newton-fp64").cpu().numpy() @ w
    fd = ((yp - ym) / (2 * h)).reshape(P, d)
    v = torch.as_tensor(np.tile(w, (P, 1)), device="cuda")
    g = ip.vjp(torch.as_tensor(x, device="cuda"), v, precision="newton-fp64").cpu().numpy()
    e = err(g, fd)
    print(f"{name} adjoint vs central differences (h = 1e-6): {e:.2e}", flush=True)
    assert e <= 1e-6

"""GPU: GEMM-shaped evaluation (large d_out) and the hand-written adjoint.

References are the oracle's extended-precision values (the exact interpolant of the FP64 data,
tests/test_oracle_extended.py).  Tolerances (max |diff| / max |ref|): the Newton form in FP64
1e-13 (no cancellation between its terms); the barycentric combination technique in FP64 within
4x the FP64 oracle's own distance from the exact values (the rounding that the Smolyak
coefficients amplify, the reference's as well), at least 1e-14; mixed (FP64 head rows, FP32 MFMA
with FP64 accumulation) 1e-6; bf16 (FP64 head rows, BF16 MFMA) 3e-3 (measured 2e-4 to 8e-4); gradient and adjoint in
FP64 1e-12, mixed adjoint 1e-6; central finite differences with h = 1e-6 in FP64: 1e-6.
"""

import numpy as np
import pytest
import torch

from oracle import indices as OI
from oracle import nodes as ON
from oracle import smolyak as OS

pytestmark = pytest.mark.gpu


NEWTON_TOL = 1e-13
MIXED_TOL = 1e-6
BF16_TOL = 3e-3
GRAD_TOL = 1e-12


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def exact(ref, x):
    return np.asarray(ref(x, dtype=np.longdouble), dtype=np.float64)


def exact_grad(ref, x):
    return np.asarray(ref.gradient(x, dtype=np.longdouble), dtype=np.float64)


def check_paths(ip, ref, x):
    ye = exact(ref, x)
    tol64 = max(4 * rel(ref(x), ye), 1e-14)
    errs = {p: rel(ip.evaluate(x, p).cpu().numpy(), ye)
            for p in ("fp64", "newton-fp64", "mixed", "bf16")}
    print("errors from the exact interpolant:", {k: f"{v:.1e}" for k, v in errs.items()},
          f"(fp64 bound {tol64:.1e})", flush=True)
    assert errs["fp64"] <= tol64
    assert errs["newton-fp64"] <= NEWTON_TOL
    assert errs["mixed"] <= MIXED_TOL
    assert errs["bf16"] <= BF16_TOL


def make(d, d_out, m, seed=0, kind="leja"):
    from smolyax_amd import nodes as SN
    from smolyax_amd.interpolation import SmolyakBarycentricInterpolator

    rng = np.random.default_rng(seed)
    k = np.log((2 + np.arange(d)) / np.log(2))
    ong, sng = (ON.leja(dim=d), SN.Leja(dim=d)) if kind == "leja" else (ON.gauss_hermite(dim=d), SN.GaussHermite(dim=d))
    t = OI.find_approximate_threshold(k, m, ong.is_nested)
    A = rng.standard_normal((d_out, d)) / (np.arange(d) + 1.0) ** 1.5
    b = rng.uniform(0, 2 * np.pi, d_out)
    f = lambda x: np.sin(np.asarray(x) @ A.T + b)
    ref = OS.Interpolant(d_out=d_out, node_gen=ong, k=k, t=t, f=f)
    ip = SmolyakBarycentricInterpolator(d_out=d_out, node_gen=sng, k=k, t=t)
    ip.set_f(f=f, vectorized=True)
    return ref, ip, rng


@pytest.mark.parametrize("d,d_out,m", [(6, 32, 200), (20, 64, 800), (50, 40, 1500)])
def test_gemm_paths(gpu, d, d_out, m):
    ref, ip, rng = make(d, d_out, m, seed=d)
    x = rng.uniform(-1, 1, (200, d))
    check_paths(ip, ref, x)


def test_gemm_paths_with_factor_table_beyond_lds(gpu):
    """d_in = 1000: the tile's factor table (764 rows) exceeds LDS in FP64 and in FP32, so the
    Newton basis computes each factor value from the coordinate where a node needs it; the GEMM
    paths still match the oracle."""
    ref, ip, rng = make(1000, 3, 2500, seed=9)
    assert ip._nlayout.n_tab_rows * 64 * 4 > 160 * 1024  # beyond LDS in FP32 (and FP64)
    x = rng.uniform(-1, 1, (300, 1000))
    check_paths(ip, ref, x)


def test_newton_plan_with_node_table_beyond_lds(gpu):
    """d_in = 6, |Lambda| ~ 4000: 1122 internal nodes, so the streaming kernels' node table does
    not fit in LDS; the plan is kept and __call__ (d_out = 1) takes the basis + contraction, the
    Jacobian and the adjoint their usual kernels; all match the oracle."""
    from smolyax_amd import _native as N

    ref, ip, rng = make(6, 1, 4000, seed=12)
    assert ip._nplan is not None and ip._nlayout.n_internal + 1 > 640
    x = rng.uniform(-1, 1, (200, 6))
    y = torch.empty((200, 1), dtype=torch.float64, device="cuda")
    xt = torch.as_tensor(x, device="cuda")
    rc = N.fn["sbi_newton_eval_mixed"](ip._nplan, xt.data_ptr(), 200, 6, y.data_ptr(), None)
    assert rc == N.SBI_ERR_CAPACITY
    ye = exact(ref, x)
    assert rel(ip.evaluate(x, "newton-fp64").cpu().numpy(), ye) <= NEWTON_TOL
    assert rel(ip.evaluate(x, "mixed").cpu().numpy(), ye) <= MIXED_TOL
    Je = exact_grad(ref, x)
    assert rel(ip.gradient(x).cpu().numpy(), Je) <= GRAD_TOL
    v = rng.standard_normal((200, 1))
    gr = np.einsum("pm,pmj->pj", v, Je)
    assert rel(ip.vjp(x, v, "fp64").cpu().numpy(), gr) <= GRAD_TOL


def test_gemm_point_chunks(gpu):
    """Chunking over points (memory_limit) gives the same result."""
    ref, ip, rng = make(8, 24, 300, seed=3)
    x = rng.uniform(-1, 1, (500, 8))
    y_all = ip.evaluate(x, "newton-fp64").cpu().numpy()
    ip._memory_limit = 1e-5  # forces chunks of 64 points
    y_ch = ip.evaluate(x, "newton-fp64").cpu().numpy()
    y_b = ip.evaluate(x, "fp64").cpu().numpy()
    assert np.array_equal(y_all, y_ch)
    ye = exact(ref, x)
    assert rel(y_b, ye) <= max(4 * rel(ref(x), ye), 1e-14)


def test_point_chunks_of_the_adjoint_and_jacobian(gpu):
    """The adjoint and the Jacobian in chunks of 64 points (memory_limit) equal one call bit for
    bit: each point's adjoint and Jacobian are computed from its own row of x and column of G^T
    only, and sbi_contract forms each element of G^T from its own row of S and of v in a fixed
    order."""
    ref, ip, rng = make(10, 12, 400, seed=4)
    x = rng.uniform(-1, 1, (300, 10))
    v = rng.standard_normal((300, 12))
    g_all = ip.vjp(x, v, "fp64").cpu().numpy()
    J_all = ip.gradient(x).cpu().numpy()
    ip._memory_limit = 1e-6  # chunks of 64 points
    g_ch = ip.vjp(x, v, "fp64").cpu().numpy()
    J_ch = ip.gradient(x).cpu().numpy()
    assert np.array_equal(g_ch, g_all)
    assert np.array_equal(J_ch, J_all)
    gr = np.einsum("pm,pmj->pj", v, exact_grad(ref, x))
    assert rel(g_ch, gr) <= GRAD_TOL


@pytest.mark.parametrize("d,d_out,m", [(3, 2, 40), (12, 20, 600), (40, 16, 1500)])
def test_vjp_matches_oracle_jacobian(gpu, d, d_out, m):
    ref, ip, rng = make(d, d_out, m, seed=10 + d)
    x = rng.uniform(-0.95, 0.95, (64, d))
    v = rng.standard_normal((64, d_out))
    g_ref = np.einsum("pm,pmj->pj", v, exact_grad(ref, x))
    g64 = ip.vjp(x, v, "fp64").cpu().numpy()
    g32 = ip.vjp(x, v, "mixed").cpu().numpy()
    print(f"adjoint: fp64 {rel(g64, g_ref):.1e}, mixed {rel(g32, g_ref):.1e}", flush=True)
    assert rel(g64, g_ref) <= GRAD_TOL
    assert rel(g32, g_ref) <= MIXED_TOL


def test_vjp_finite_differences(gpu):
    """Adjoint of w.y against central differences with h = 1e-6 in FP64."""
    d, d_out = 10, 8
    ref, ip, rng = make(d, d_out, 500, seed=4)
    x = rng.uniform(-0.9, 0.9, (16, d))
    w = rng.standard_normal(d_out)
    g = ip.vjp(x, np.tile(w, (16, 1)), "fp64").cpu().numpy()
    h = 1e-6
    fd = np.zeros_like(g)
    for j in range(d):
        e = np.zeros(d)
        e[j] = h
        yp = ip.evaluate(x + e, "newton-fp64").cpu().numpy() @ w
        ym = ip.evaluate(x - e, "newton-fp64").cpu().numpy() @ w
        fd[:, j] = (yp - ym) / (2 * h)
    assert rel(g, fd) < 1e-6


def test_autograd_function(gpu):
    d, d_out = 8, 4
    ref, ip, rng = make(d, d_out, 300, seed=5)
    x = torch.tensor(rng.uniform(-0.9, 0.9, (32, d)), device="cuda", requires_grad=True)
    y = ip.differentiable(x, "newton-fp64")
    w = torch.tensor(rng.standard_normal((32, d_out)), device="cuda")
    (y * w).sum().backward()
    g_ref = np.einsum("pm,pmj->pj", w.cpu().numpy(), exact_grad(ref, x.detach().cpu().numpy()))
    assert rel(x.grad.cpu().numpy(), g_ref) <= GRAD_TOL


def test_vjp_non_nested_matches_oracle(gpu):
    """Gauss-Hermite nodes: the adjoint of the Newton form on the proxy grid (same polynomial)."""
    ref, ip, rng = make(4, 3, 60, seed=6, kind="gh")
    x = rng.standard_normal((20, 4)) / np.sqrt(2)
    v = rng.standard_normal((20, 3))
    g_ref = np.einsum("pm,pmj->pj", v, exact_grad(ref, x))
    assert rel(ip.vjp(x, v).cpu().numpy(), g_ref) <= GRAD_TOL


def test_jacobian_many_outputs_matches_adjoint(gpu):
    """Dense Jacobian at a size the oracle does not reach (700 points, d_out 600: several point
    tiles and several passes of outputs per workgroup) against the FP64 adjoint, which shares no
    kernel with it: v^T J == vjp(x, v) to 1e-10 relative."""
    d, d_out = 20, 600
    ref, ip, rng = make(d, d_out, 400, seed=8)
    x = rng.uniform(-1, 1, (700, d))
    v = rng.standard_normal((700, d_out))
    J = ip.gradient(x)
    assert tuple(J.shape) == (700, d_out, d)
    g = np.einsum("pm,pmj->pj", v, J.cpu().numpy())
    assert rel(g, ip.vjp(x, v, "fp64").cpu().numpy()) <= GRAD_TOL
    # the first points against the exact Jacobian
    assert rel(J[:3].cpu().numpy(), exact_grad(ref, x[:3])) <= GRAD_TOL


def test_vjp_global_factor_table(gpu):
    """400 input dims: the factor table (369 rows) exceeds LDS and is read from the global table;
    dims with more than 128 adjoint entries are split into units whose partial sums are added in
    a fixed order (FP64 and FP32 G)."""
    ref, ip, rng = make(400, 3, 1000, seed=12)
    assert ip._nlayout.n_tab_rows * 64 * 8 > 160 * 1024 and ip._nlayout.n_jac_red > 0
    x = rng.uniform(-1, 1, (70, 400))
    v = rng.standard_normal((70, 3))
    g_ref = np.einsum("pm,pmj->pj", v, exact_grad(ref, x))
    assert rel(ip.vjp(x, v, "fp64").cpu().numpy(), g_ref) <= GRAD_TOL
    assert rel(ip.vjp(x, v, "mixed").cpu().numpy(), g_ref) <= MIXED_TOL


@pytest.mark.parametrize("d,d_out,m,P", [(8, 37, 400, 1), (8, 37, 400, 100), (6, 300, 900, 700),
                                         (12, 5, 1500, 5000), (30, 64, 2000, 256)])
def test_mixed_gemm_one_call_equals_steps(gpu, d, d_out, m, P):
    """sbi_newton_eval_mixed_gemm (head rows, FP32 rows and the contraction in one call, k split
    over more workgroups for small batches) gives the bits of its three steps called one by one
    with the unsplit contraction (sbi_newton_rows_f64, sbi_newton_basis_rows dtype 3,
    sbi_contract_split_k4)."""
    import ctypes

    from smolyax_amd import _native as N

    print(f"mixed one call d={d} d_out={d_out} m={m} P={P}", flush=True)
    _, ip, _ = make(d, d_out, m, seed=d + P)
    x = torch.as_tensor(np.random.default_rng(P).uniform(-1, 1, (P, d)), device=gpu)
    y1 = ip.evaluate(x, "mixed")
    nd = ip._ndesc
    nh, KQ, Nn = nd.gemm_head, nd.gemm_kq, nd.basis_n
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    HT = torch.empty((max(KQ // 4, 1), P, 4), dtype=torch.float32, device=gpu)
    H64 = torch.empty((max(nh, 1), P), dtype=torch.float64, device=gpu)
    if Nn > nh:
        N.check(N.fn["sbi_newton_basis_rows"](ip._nplan, x.data_ptr(), P, x.stride(0), nh, Nn,
                                              HT.data_ptr(), P, 3, st), "basis")
    if nh:
        N.check(N.fn["sbi_newton_rows_f64"](ip._nplan, x.data_ptr(), P, x.stride(0),
                                            nd.head_fac, nh, H64.data_ptr(), P, st), "rows")
    y0 = torch.full((P, d_out), float("nan"), dtype=torch.float64, device=gpu)
    N.check(N.fn["sbi_contract_split_k4"](H64.data_ptr(), P, nd.gemm_s64, nd.ld_s64, nh,
                                          HT.data_ptr() if KQ else None, P,
                                          nd.gemm_sq if KQ else None, d_out, KQ, None,
                                          y0.data_ptr(), d_out, P, d_out, st), "contract")
    torch.cuda.synchronize()
    assert torch.equal(y0, y1)

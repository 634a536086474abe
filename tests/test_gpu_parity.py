"""GPU parity: the HIP kernels (through the C ABI) against the FP64 oracle.

Tolerances: FP64 evaluation, gradient and integral agree with the oracle to <= 1e-12 relative
(max |diff| / max |y|).  The oracle follows the reference's formula; the kernels use the first
(modified Lagrange) form, so agreement is to rounding, amplified by the Smolyak coefficients.
"""

import numpy as np
import pytest

from oracle import indices as OI
from oracle import nodes as ON
from oracle import smolyak as OS

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def build_pair(d, d_out, m, f, kind="leja", seed=0, custom=False):
    from smolyax_amd import nodes as SN
    from smolyax_amd.interpolation import SmolyakBarycentricInterpolator

    rng = np.random.default_rng(seed)
    k = np.log((2 + np.arange(d)) / np.log(2))
    if kind == "leja":
        if custom:
            dom = np.sort(rng.random((d, 2)), axis=1)
            dom[:, 1] += 0.1
            ong, sng = ON.leja(domains=dom), SN.Leja(domains=dom)
        else:
            ong, sng = ON.leja(dim=d), SN.Leja(dim=d)
    else:
        if custom:
            mean, sc = rng.standard_normal(d), rng.random(d) + 0.3
            ong, sng = ON.gauss_hermite(mean=mean, scaling=sc), SN.GaussHermite(mean, sc)
        else:
            ong, sng = ON.gauss_hermite(dim=d), SN.GaussHermite(dim=d)
    t = OI.find_approximate_threshold(k, m, ong.is_nested)
    ref = OS.Interpolant(d_out=d_out, node_gen=ong, k=k, t=t, f=f)
    ip = SmolyakBarycentricInterpolator(d_out=d_out, node_gen=sng, k=k, t=t, f=f)
    return ref, ip, ong


def sample(ong, P, seed):
    rng = np.random.default_rng(seed)
    if ong.is_nested:
        x = rng.uniform(-1, 1, (P, ong.dim))
    else:
        x = rng.standard_normal((P, ong.dim)) / np.sqrt(2)
    return np.stack([g.scale(x[:, j]) for j, g in enumerate(ong)], axis=1)


CASES = [
    # d, d_out, m, kind, custom
    (1, 1, 9, "leja", False),
    (3, 2, 60, "leja", True),
    (5, 3, 150, "gh", False),
    (4, 5, 120, "gh", True),
    (10, 1, 1000, "leja", False),
    (8, 9, 400, "leja", False),
    (30, 2, 600, "leja", False),
]


@pytest.mark.parametrize("d,d_out,m,kind,custom", CASES)
def test_eval_parity(gpu, d, d_out, m, kind, custom):
    rng = np.random.default_rng(d * 100 + d_out)
    A = rng.standard_normal((d_out, d)) / (np.arange(d) + 1.0)
    f = lambda x: np.cos(A @ x + 0.2) + 0.1 * (A @ x) ** 2
    ref, ip, ong = build_pair(d, d_out, m, f, kind, seed=d, custom=custom)
    x = sample(ong, 333, seed=7)
    y = ip(x).cpu().numpy()
    yr = ref(x)
    assert y.shape == yr.shape == (333, d_out)
    assert rel(y, yr) <= TOL, rel(y, yr)
    # single point (d_in,) input promoted to (1, d_in)
    y1 = ip(x[5]).cpu().numpy()
    assert y1.shape == (1, d_out)
    assert rel(y1, yr[5:6]) <= TOL


def test_eval_at_grid_nodes(gpu):
    """One-hot branch: at sparse-grid nodes the interpolant reproduces f."""
    d, d_out = 4, 2
    f = lambda x: np.array([np.exp(np.sum(x) / 3), np.sin(x[0] - x[3])])
    ref, ip, ong = build_pair(d, d_out, 80, f)
    k = np.log((2 + np.arange(d)) / np.log(2))
    t = OI.find_approximate_threshold(k, 80, True)
    lam = OI.indexset(k, t)
    xi = ong[0](12)
    x = np.array([[xi[dict(nu).get(j, 0)] for j in range(d)] for nu in lam])
    y = ip(x).cpu().numpy()
    assert rel(y, ref(x)) <= TOL
    fx = np.array([f(r) for r in x])
    assert np.max(np.abs(y - fx)) < 1e-12


@pytest.mark.parametrize("d,d_out,m,kind,custom", CASES[:6])
def test_gradient_parity(gpu, d, d_out, m, kind, custom):
    rng = np.random.default_rng(d)
    A = rng.standard_normal((d_out, d))
    f = lambda x: np.tanh(A @ x / d)
    ref, ip, ong = build_pair(d, d_out, m, f, kind, seed=d + 1, custom=custom)
    x = sample(ong, 70, seed=11)
    J = ip.gradient(x).cpu().numpy()
    Jr = ref.gradient(x)
    assert J.shape == Jr.shape == (70, d_out, d)
    assert rel(J, Jr) <= 1e-11, rel(J, Jr)


def test_gradient_nan_at_nodes(gpu):
    d = 3
    f = lambda x: np.array([x[0] * x[1] + x[2] ** 2])
    ref, ip, ong = build_pair(d, 1, 30, f)
    x = np.array([[0.3, 0.0, -0.2], [0.25, 0.5, 0.125]])
    J = ip.gradient(x).cpu().numpy()
    Jr = ref.gradient(x)
    assert np.array_equal(np.isnan(J), np.isnan(Jr))
    assert np.isnan(J[0, 0, 1]) and not np.isnan(J[1]).any()
    ok = ~np.isnan(Jr)
    assert rel(J[ok], Jr[ok]) <= 1e-12


def test_gradient_nan_at_gauss_hermite_nodes(gpu):
    """Non-nested nodes (Newton form on the Leja proxy grid): NaN exactly where x_j is a node of a
    Gauss-Hermite degree that some block uses in dim j; equal to the oracle elsewhere."""
    d, d_out = 3, 2
    rng = np.random.default_rng(5)
    A = rng.standard_normal((d_out, d))
    f = lambda x: np.sin(A @ x)
    ref, ip, ong = build_pair(d, d_out, 60, f, kind="gh")
    x = sample(ong, 30, seed=3)
    p = 0
    for j in range(d):
        for deg in range(0, 6):
            x[p, j] = ong[j](deg)[deg // 2]
            p += 1
    J = ip.gradient(x).cpu().numpy()
    Jr = ref.gradient(x)
    assert np.array_equal(np.isnan(J), np.isnan(Jr))
    assert np.isnan(Jr).any() and not np.isnan(Jr).all()
    ok = ~np.isnan(Jr)
    assert rel(J[ok], Jr[ok]) <= 1e-11, rel(J[ok], Jr[ok])


@pytest.mark.parametrize("d_out", [1, 12])
def test_gradient_global_factor_table(gpu, d_out):
    """400 input dims: the Newton factor table (369 rows) exceeds LDS and is kept in global
    memory; both the register path (d_out <= 8) and the matrix-core path; NaN at a node."""
    d = 400
    rng = np.random.default_rng(4)
    A = rng.standard_normal((d_out, d)) / (np.arange(d) + 1.0)
    f = lambda x: np.sin(A @ x + 0.4)
    ref, ip, ong = build_pair(d, d_out, 1000, f)
    x = sample(ong, 5, seed=2)
    x[1, 0] = ong[0](3)[2]
    J = ip.gradient(x).cpu().numpy()
    Jr = ref.gradient(x)
    assert np.array_equal(np.isnan(J), np.isnan(Jr)) and np.isnan(J[1, :, 0]).all()
    ok = ~np.isnan(Jr)
    assert rel(J[ok], Jr[ok]) <= 1e-11, rel(J[ok], Jr[ok])


def test_bary_jacobian_kernel(gpu):
    """The barycentric Jacobian kernel (sbi_bary_jacobian_f64, used when the Newton form is
    unavailable) through the C ABI: two output passes, two point tiles, NaN at a node."""
    import ctypes

    import torch

    from smolyax_amd import _native as N

    d, d_out = 6, 11
    rng = np.random.default_rng(9)
    A = rng.standard_normal((d_out, d))
    f = lambda x: np.tanh(A @ x / d)
    ref, ip, ong = build_pair(d, d_out, 200, f)
    # coordinates kept 0.01 away from the end nodes +-1: at 2.7e-6 from a node the reference's
    # formula (b / sum b with |b| ~ 1 / |x - xi|) loses 4.8e-10 relative against exact rational
    # arithmetic while the first form used here stays within 4e-14
    x = np.clip(sample(ong, 90, seed=5), -0.99, 0.99)
    x[7, 2] = ong[2](4)[1]
    xt = torch.as_tensor(x, device="cuda")
    J = torch.empty((90, d_out, d), dtype=torch.float64, device="cuda")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    N.check(N.fn["sbi_bary_jacobian_f64"](ip._plan, xt.data_ptr(), 90, xt.stride(0), J.data_ptr(),
                                          st), "jacobian")
    J = J.cpu().numpy()
    Jr = ref.gradient(x)
    assert np.array_equal(np.isnan(J), np.isnan(Jr)) and np.isnan(J[7, :, 2]).all()
    ok = ~np.isnan(Jr)
    assert rel(J[ok], Jr[ok]) <= 1e-11


@pytest.mark.parametrize("d,d_out,m,kind,custom", CASES)
def test_integral_parity(gpu, d, d_out, m, kind, custom):
    rng = np.random.default_rng(d + 5)
    A = rng.standard_normal((d_out, d))
    f = lambda x: np.exp(-0.3 * (A @ x) ** 2)
    ref, ip, ong = build_pair(d, d_out, m, f, kind, seed=d + 2, custom=custom)
    Q = ip.integral().cpu().numpy()
    Qr = ref.integral()
    assert Q.shape == (d_out,)
    assert rel(Q, Qr) <= TOL, rel(Q, Qr)


def test_config1_parity(gpu):
    """Plain FP64 CPU-runnable workload: d_in=10, |Lambda|~1000, 1024 points from seed 0."""
    from smolyax_amd import workloads as W

    d = 10
    f = W.f_rational(d)
    ref, ip, ong = build_pair(d, 1, 1000, f)
    x = W.points_uniform(1024, d, seed=0)
    assert rel(ip(x).cpu().numpy(), ref(x)) <= TOL


def test_f_evals_reuse(gpu):
    """set_f returns / accepts the reference's f_evals dict; reuse makes no new calls."""
    from smolyax_amd import nodes as SN
    from smolyax_amd.interpolation import SmolyakBarycentricInterpolator

    d = 5
    k = np.log((2 + np.arange(d)) / np.log(2))
    calls = []
    f = lambda x: calls.append(1) or np.sum(x ** 2)
    for ng in (SN.Leja(dim=d), SN.GaussHermite(dim=d)):
        t = OI.find_approximate_threshold(k, 100, ng.is_nested)
        ip = SmolyakBarycentricInterpolator(d_out=1, node_gen=ng, k=k, t=t)
        calls.clear()
        fe = ip.set_f(f=f)
        assert len(calls) == ip.n_f_evals_new == ip.n_f_evals
        ip2 = SmolyakBarycentricInterpolator(d_out=1, node_gen=ng, k=k, t=t)
        calls.clear()
        ip2.set_f(f=f, f_evals=fe)
        assert len(calls) == 0 and ip2.n_f_evals_new == 0
        x = np.random.default_rng(0).uniform(-0.5, 0.5, (9, d))
        assert rel(ip(x).cpu().numpy(), ip2(x).cpu().numpy()) == 0.0


def test_input_shapes_and_errors(gpu):
    """interpolation.py:253-262: (d_in,) is one point; empty batches give empty results; a wrong
    d_in and use before set_f raise AssertionError.  Every precision path on the same inputs."""
    from smolyax_amd import nodes as SN
    from smolyax_amd.interpolation import SmolyakBarycentricInterpolator

    d, d_out = 4, 3
    f = lambda x: np.array([np.sum(x), np.prod(1 + x), np.exp(x[0])])
    ref, ip, ong = build_pair(d, d_out, 80, f)
    x1 = np.array([0.1, -0.3, 0.7, 0.2])
    for prec in ("fp64", "mixed", "newton-fp64", "bf16", "bary-fp32"):
        y0 = ip.evaluate(np.zeros((0, d)), prec)
        assert tuple(y0.shape) == (0, d_out)
        y1 = ip.evaluate(x1, prec).cpu().numpy()
        assert y1.shape == (1, d_out)
        assert rel(y1, ref(x1[None, :])) <= {"fp64": TOL, "newton-fp64": 1e-12}.get(prec, 1e-2)
    assert tuple(ip.gradient(np.zeros((0, d))).shape) == (0, d_out, d)
    assert tuple(ip.gradient(x1).shape) == (1, d_out, d)
    with pytest.raises(AssertionError):
        ip(np.zeros((2, d + 1)))
    fresh = SmolyakBarycentricInterpolator(d_out=1, node_gen=SN.Leja(dim=d), k=[1.0] * d, t=2.0)
    with pytest.raises(AssertionError):
        fresh(x1)
    with pytest.raises(AssertionError):
        fresh.integral()


@pytest.mark.parametrize("d,m", [(1, 22), (1, 31), (2, 90)])
def test_gauss_hermite_high_degree(gpu, d, m):
    """Gauss-Hermite interpolants of degree 20 to 30: eval, gradient and adjoint match the oracle
    whichever form serves them (the proxy Newton form where it meets FP64 accuracy on sample
    points, the barycentric kernels otherwise); the proxy's measured error is reported."""
    A = np.random.default_rng(d + m).standard_normal((2, d)) * 0.3
    f = lambda x: np.sin(A @ np.asarray(x) + 0.4)
    ref, ip, ong = build_pair(d, 2, m, f, kind="gh")
    x = sample(ong, 150, seed=m)
    print(f"GH d={d} m={m}: proxy error {ip.proxy_error}, mixed path {ip.mixed_path}", flush=True)
    assert rel(ip.evaluate(x, "fp64").cpu().numpy(), ref(x)) < TOL
    assert rel(ip.gradient(x).cpu().numpy(), ref.gradient(x)) < 1e-11
    v = np.random.default_rng(1).standard_normal((len(x), 2))
    g = ip.vjp(x, v, precision="fp64").cpu().numpy()
    assert rel(g, np.einsum("pm,pmj->pj", v, ref.gradient(x))) < 1e-11
    assert rel(ip.evaluate(x, "mixed").cpu().numpy(), ref(x)) < 1e-6 or ip.mixed_path != "newton"
    if ip.proxy_error is not None and ip.proxy_error > 1e-12:
        with pytest.raises(ValueError):
            ip.evaluate(x, "newton-fp64")


def test_mixed_without_newton_form_warns(gpu):
    """An index set past the Newton layout's limits (nodes with 8 factors: the adjoint records
    hold 6 other factors) has no Newton form: 'mixed' is served by 'bary-fp32' with a
    SmolyaxPrecisionWarning, and the Newton-only precisions raise."""
    from smolyax_amd import nodes as SN
    from smolyax_amd.interpolation import SmolyakBarycentricInterpolator, SmolyaxPrecisionWarning

    d = 8
    k = np.ones(d)
    t = 8.5
    f = lambda x: np.exp(-np.sum(np.asarray(x) / (1.0 + np.arange(d))))
    ip = SmolyakBarycentricInterpolator(d_out=1, node_gen=SN.Leja(dim=d), k=k, t=t, f=f,
                                        precision="mixed")
    assert ip.mixed_path == "bary-fp32"
    x = np.random.default_rng(0).uniform(-1, 1, (64, d))
    with pytest.warns(SmolyaxPrecisionWarning):
        y = ip(x)
    y32 = ip.evaluate(x, "bary-fp32")
    assert np.array_equal(y.cpu().numpy(), y32.cpu().numpy())
    y64 = ip.evaluate(x, "fp64").cpu().numpy()
    e = rel(y.cpu().numpy(), y64)
    print(f"bary-fp32 in place of mixed: {e:.2e} from fp64", flush=True)
    assert e < 1e-3
    for prec in ("newton-fp64", "bf16"):
        with pytest.raises(ValueError):
            ip.evaluate(x, prec)


@pytest.mark.parametrize("d,m,kind,kscale", [(5, 300, "gh", 1.0), (8, 3000, "leja", 0.3),
                                             (2, 40, "leja", 1.0), (12, 600, "leja", 0.2),
                                             (1, 22, "gh", 1.0), (6, 200, "leja", 1.0)])
def test_eval_record_program(gpu, d, m, kind, kscale):
    """d_out = 1 FP64 evaluation runs the record-program kernel: against the oracle, and against
    the per-block kernel (the same plan without its program).  The index sets cover blocks with
    four or more active dims and streamed sizes outside 2..4 (the kernel's per-block path), odd
    prefix degrees (records padded with zeros), points on grid nodes and a partial last tile."""
    import ctypes

    import torch

    from smolyax_amd import _native as N
    from smolyax_amd import nodes as SN
    from smolyax_amd.interpolation import SmolyakBarycentricInterpolator

    rng = np.random.default_rng(d + m)
    k = kscale * np.log((2 + np.arange(d)) / np.log(2))
    ong, sng = (ON.leja(dim=d), SN.Leja(dim=d)) if kind == "leja" else \
        (ON.gauss_hermite(dim=d), SN.GaussHermite(dim=d))
    t = OI.find_approximate_threshold(k, m, ong.is_nested)
    A = rng.standard_normal((1, d)) / (np.arange(d) + 1.0)
    f = lambda x: np.cos(A @ x + 0.2) + 0.1 * (A @ x) ** 2
    ref = OS.Interpolant(d_out=1, node_gen=ong, k=k, t=t, f=f)
    ip = SmolyakBarycentricInterpolator(d_out=1, node_gen=sng, k=k, t=t, f=f)
    assert ip._prog is not None and ip._desc.n_pseg > 0
    x = sample(ong, 203, seed=m)
    x[:3, 0] = ong[0](4)[:3]                 # points on nodes of dim 0
    print(f"record program {kind} d={d} m={m}: {ip._prog.n_records} records", flush=True)
    y = ip.evaluate(x, "fp64").cpu().numpy()
    # against the exact interpolant of the FP64 data (the oracle in extended precision): the
    # larger index sets carry sum |zeta| |F| of 1e2..1e3 max |y|, so every FP64 evaluation of the
    # combination technique, the oracle's included, rounds at 1e-13..1e-12 there
    y_ex = ref(x, dtype=np.longdouble).astype(np.float64)
    e_ora = rel(ref(x), y_ex)
    e_prog = rel(y, y_ex)
    print(f"  record program {e_prog:.2e}, FP64 oracle {e_ora:.2e} from the exact values",
          flush=True)
    assert e_prog <= max(4.0 * e_ora, 1e-14)
    # the per-block kernel on the same device tables
    dA = N.BaryDesc()
    ctypes.pointer(dA)[0] = ip._desc
    dA.pseg = None
    plan = ctypes.c_void_p()
    N.check(N.fn["sbi_bary_plan_create"](ctypes.byref(dA), ctypes.byref(plan)), "plan")
    try:
        xt = torch.as_tensor(x, device="cuda")
        y0 = torch.empty((len(x), 1), dtype=torch.float64, device="cuda")
        N.check(N.fn["sbi_bary_eval_f64"](plan, xt.data_ptr(), len(x), d, y0.data_ptr(), None),
                "eval")
        torch.cuda.synchronize()
    finally:
        N.fn["sbi_plan_destroy"](plan)
    e_blk = rel(y0.cpu().numpy(), y_ex)
    print(f"  per-block kernel {e_blk:.2e}", flush=True)
    assert e_blk <= max(4.0 * e_ora, 1e-14)

"""GPU: the hierarchical (Newton) kernels against the FP64 oracle.

References: the oracle's extended-precision values (the exact interpolant of the FP64 data).
newton-fp64: 1e-13 relative (max |diff| / max |y|; the Newton form has no cancelling terms);
the barycentric FP64 path within 4x the FP64 oracle's own distance from the exact values (at least
1e-14); mixed: 1e-6, the precision target of the mixed path.
"""

import numpy as np
import pytest
import torch

from oracle import indices as OI
from oracle import nodes as ON
from oracle import smolyak as OS

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


@pytest.mark.parametrize("d,m,dom", [(1, 9, False), (4, 60, True), (10, 1000, False),
                                     (40, 2000, False), (200, 5000, False)])
def test_newton_paths(gpu, d, m, dom):
    from smolyax_amd import nodes as SN
    from smolyax_amd.interpolation import SmolyakBarycentricInterpolator

    rng = np.random.default_rng(d)
    k = np.log((2 + np.arange(d)) / np.log(2))
    if dom:
        dm = np.sort(rng.random((d, 2)), axis=1) + [0, 0.3]
        ong, sng = ON.leja(domains=dm), SN.Leja(domains=dm)
    else:
        ong, sng = ON.leja(dim=d), SN.Leja(dim=d)
    t = OI.find_approximate_threshold(k, m, True)
    a = rng.standard_normal(d) / (np.arange(d) + 1.0) ** 2
    f = lambda x: 1.0 / (1.5 + np.tanh(np.asarray(x) @ a))
    ref = OS.Interpolant(d_out=1, node_gen=ong, k=k, t=t, f=f)
    ip = SmolyakBarycentricInterpolator(d_out=1, node_gen=sng, k=k, t=t, f=f, precision="mixed")
    P = 300
    x = np.stack([g.scale(c) for g, c in zip(ong, rng.uniform(-1, 1, (P, d)).T)], axis=1)
    ye = np.asarray(ref(x, dtype=np.longdouble), dtype=np.float64)
    e_o64 = rel(ref(x), ye)
    e_mixed = rel(ip(x).cpu().numpy(), ye)
    e_n64 = rel(ip.evaluate(x, precision="newton-fp64").cpu().numpy(), ye)
    e_b64 = rel(ip.evaluate(x, precision="fp64").cpu().numpy(), ye)
    print(f"d={d} m={m}: newton-fp64 {e_n64:.1e} fp64 {e_b64:.1e} (oracle FP64 {e_o64:.1e}) "
          f"mixed {e_mixed:.1e}", flush=True)
    assert e_n64 <= 1e-13, e_n64
    assert e_b64 <= max(4 * e_o64, 1e-14), (e_b64, e_o64)
    assert e_mixed <= 1e-6, e_mixed


def test_newton_form_of_non_nested_interpolant(gpu):
    """Gauss-Hermite interpolants get a Newton form on a nested proxy grid (the same polynomial):
    its FP64 evaluation matches the barycentric FP64 path, the mixed paths are FP32-accurate."""
    from smolyax_amd import nodes as SN
    from smolyax_amd.interpolation import SmolyakBarycentricInterpolator

    for d, d_out, m in [(2, 1, 30), (6, 1, 300), (5, 20, 400)]:
        k = np.log((2 + np.arange(d)) / np.log(2))
        t = OI.find_approximate_threshold(k, m, False)
        rng = np.random.default_rng(d)
        a = rng.standard_normal((d_out, d)) / (1.0 + np.arange(d))
        f = lambda x: np.sin(a @ x + 0.3)
        ip = SmolyakBarycentricInterpolator(d_out=d_out, node_gen=SN.GaussHermite(dim=d), k=k, t=t,
                                            f=f, precision="mixed")
        assert ip._nplan is not None
        x = torch.randn((300, d), dtype=torch.float64, device="cuda")
        ref = ip.evaluate(x, "fp64")
        scale = float(ref.abs().max())
        print(f"GH d={d} m={m}: proxy error {ip.proxy_error:.1e}", flush=True)
        if ip.proxy_error <= 1e-12:
            assert float((ip.evaluate(x, "newton-fp64") - ref).abs().max()) <= 1e-11 * scale
        else:
            with pytest.raises(ValueError):
                ip.evaluate(x, "newton-fp64")
        assert float((ip.evaluate(x, "mixed") - ref).abs().max()) <= 1e-6 * scale
        assert float((ip(x) - ref).abs().max()) <= 1e-6 * scale


@pytest.mark.parametrize("d", [1, 5, 6, 40, 41])
@pytest.mark.parametrize("extra", [1, 2])
def test_mixed_strided_points_same_bits(gpu, d, extra):
    """x as a row view of a wider array (row stride d + extra, odd or even) gives the same bits
    as the contiguous copy, and both match the FP64 Newton path to the mixed tolerance.  Odd d
    with an even stride is the case where a 16-byte pair load would straddle the row end."""
    from smolyax_amd import nodes as SN
    from smolyax_amd.interpolation import SmolyakBarycentricInterpolator

    k = np.log((2 + np.arange(d)) / np.log(2))
    t = OI.find_approximate_threshold(k, 300, True)
    f = lambda x: np.exp(-np.sum(x / (1 + np.arange(d))))
    ip = SmolyakBarycentricInterpolator(d_out=1, node_gen=SN.Leja(dim=d), k=k, t=t, f=f)
    big = torch.rand((300, d + extra), dtype=torch.float64, device="cuda") * 2 - 1
    xv = big[:, :d]
    assert xv.stride(0) == d + extra
    y_view = ip.evaluate(xv, "mixed")
    y_cont = ip.evaluate(xv.contiguous(), "mixed")
    assert torch.equal(y_view, y_cont)
    y64 = ip.evaluate(xv.contiguous(), "newton-fp64")
    assert float((y_view - y64).abs().max() / y64.abs().max()) < 1e-6


@pytest.mark.parametrize("d,m", [(4, 60), (10, 1000), (40, 2000), (200, 5000), (1000, 10000)])
def test_newton_fp64_matrix_cores_match_streaming(gpu, d, m):
    """sbi_newton_eval_f64 on the FP64 matrix-core kernel (W in FP64, A steps from global memory,
    v_mfma_f64_16x16x4) against the streaming FP64 kernel of the same plan (mt_a64 = NULL): the
    same Newton form in a different summation order, within 1e-14 of max |y|; a ragged last tile
    and a strided point array included."""
    import ctypes

    from smolyax_amd import _native as N
    from smolyax_amd import nodes as SN
    from smolyax_amd.interpolation import SmolyakBarycentricInterpolator

    print(f"newton fp64 matrix cores d={d} m={m}", flush=True)
    rng = np.random.default_rng(d + 1)
    k = np.log((2 + np.arange(d)) / np.log(2))
    t = OI.find_approximate_threshold(k, m, True)
    a = rng.standard_normal(d) / (np.arange(d) + 1.0) ** 2
    f = lambda x: np.exp(-(np.asarray(x) @ a))
    ip = SmolyakBarycentricInterpolator(d_out=1, node_gen=SN.Leja(dim=d), k=k, t=t, f=f)
    assert N.fn["sbi_newton_f64_matrix_cores"](ip._nplan) == 1
    desc = N.NewtonDesc()
    ctypes.memmove(ctypes.byref(desc), ctypes.byref(ip._ndesc), ctypes.sizeof(desc))
    desc.mt_a64 = None
    plan = ctypes.c_void_p()
    N.check(N.fn["sbi_newton_plan_create"](ctypes.byref(desc), ctypes.byref(plan)), "plan")
    try:
        assert N.fn["sbi_newton_f64_matrix_cores"](plan) == 0
        P = 1000
        wide = torch.as_tensor(rng.uniform(-1, 1, (P, d + 3)), device=gpu)
        for xx in (wide[:, :d].contiguous(), wide[:, 1:1 + d]):
            st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
            y_mc = torch.empty((P, 1), dtype=torch.float64, device=gpu)
            y_st = torch.empty((P, 1), dtype=torch.float64, device=gpu)
            N.check(N.fn["sbi_newton_eval_f64"](ip._nplan, xx.data_ptr(), P, xx.stride(0),
                                                y_mc.data_ptr(), st), "mc")
            N.check(N.fn["sbi_newton_eval_f64"](plan, xx.data_ptr(), P, xx.stride(0),
                                                y_st.data_ptr(), st), "stream")
            torch.cuda.synchronize()
            e = float((y_mc - y_st).abs().max() / y_st.abs().max())
            print(f"  matrix cores vs streaming: {e:.1e}", flush=True)
            assert e <= 1e-14
    finally:
        N.fn["sbi_newton_plan_destroy"](plan)

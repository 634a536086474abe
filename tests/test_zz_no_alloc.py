"""GPU: the allocation-free integral entry point (sbi_bary_integral_f64_noalloc) against the one
that takes a workspace (sbi_bary_integral_f64): the same Q bit for bit (same kernels, same
summation order), on two calls in a row of which the second has a smaller input, and again in the
opposite order (the device scratch keeps the last call's contents)."""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def build(d, d_out, m, seed):
    from smolyax_amd import indices as SI
    from smolyax_amd import nodes as SN
    from smolyax_amd.interpolation import SmolyakBarycentricInterpolator

    k = np.log((2 + np.arange(d)) / np.log(2))
    t = SI.find_approximate_threshold(k, m, True)
    A = np.random.default_rng(seed).standard_normal((d_out, d))
    ip = SmolyakBarycentricInterpolator(d_out=d_out, node_gen=SN.Leja(dim=d), k=k, t=t)
    ip.set_f(f=lambda X: np.cos(X @ A.T), vectorized=True)
    return ip


def with_workspace(ip, st):
    from smolyax_amd import _native as N

    part = torch.empty(int(N.fn["sbi_bary_integral_workspace"](ip._plan)), dtype=torch.float64,
                       device="cuda")
    Q = torch.empty(ip.d_out, dtype=torch.float64, device="cuda")
    N.check(N.fn["sbi_bary_integral_f64"](ip._plan, Q.data_ptr(), part.data_ptr(), st), "integral")
    return Q


def no_alloc(ip, st):
    from smolyax_amd import _native as N

    Q = torch.full((ip.d_out,), float("nan"), dtype=torch.float64, device="cuda")
    N.check(N.fn["sbi_bary_integral_f64_noalloc"](ip._plan, Q.data_ptr(), st), "integral_noalloc")
    return Q


def test_noalloc_integral_matches_workspace_form(gpu):
    big = build(12, 300, 2000, seed=1)
    small = build(4, 2, 60, seed=2)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for order in ([big, small], [small, big]):
        refs = [with_workspace(ip, st) for ip in order]
        torch.cuda.synchronize()
        got = [no_alloc(ip, st) for ip in order]  # two calls in a row on the shared scratch
        torch.cuda.synchronize()
        for ip, r, g in zip(order, refs, got):
            print(f"integral d_in={ip.d_in} d_out={ip.d_out}: no-alloc == workspace form",
                  flush=True)
            assert torch.equal(g, r)
    assert torch.equal(no_alloc(big, st), big.integral())


def test_noalloc_integral_rejects_oversized_plan():
    """A plan whose partial sums exceed the device scratch is refused with SBI_ERR_CAPACITY
    before any device work (the descriptor's device pointers are never touched here)."""
    from smolyax_amd import _native as N

    d = N.BaryDesc()
    d.d_in, d.d_out = 1, 1_000_000
    d.n_slots, d.dmax1, d.n_blocks, d.n_runs, d.k_raw = 1, 12, 1, 1, 1_000_000
    d.n_split = 16
    fake = 256  # non-null stand-in addresses: plan creation copies scalars only
    for name in ("F", "offset", "blk_desc", "seg_off", "run_start", "blk_fac", "blk_foff", "ent_q"):
        setattr(d, name, fake)
    plan = ctypes.c_void_p()
    assert N.fn["sbi_bary_plan_create"](ctypes.byref(d), ctypes.byref(plan)) == 0
    try:
        rc = N.fn["sbi_bary_integral_f64_noalloc"](plan, fake, None)
        assert rc == N.SBI_ERR_CAPACITY
        assert b"scratch" in N.fn["sbi_last_error"]()
    finally:
        N.fn["sbi_plan_destroy"](plan)

"""GPU: sbi_contract (Y = offset + A^T B on the matrix cores) through the C-ABI.

Reference: the same product in FP64 on the device with torch (operands upcast exactly).  Bounds,
per element, relative to sum_k |A[k,p]| |B[k,j]|: FP64 operands 1e-13 (FP64 MFMA, K <= 4096 terms);
FP32 operands 2e-6 (products and 256-term partial sums in FP32, flushed into FP64).  Rows are
independent: a point's row is identical bit for bit alone and inside a larger batch.
"""

import ctypes

import pytest
import torch

from smolyax_amd import _native as N

pytestmark = pytest.mark.gpu


def contract(A, B, offset=None, P=None, n=None):
    K = A.shape[0]
    P = A.shape[1] if P is None else P
    n = B.shape[1] if n is None else n
    Y = torch.full((P, n), float("nan"), dtype=torch.float64, device=A.device)
    code = {torch.float64: 0, torch.float32: 1}[A.dtype]
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = N.fn["sbi_contract"](A.data_ptr(), A.stride(0), B.data_ptr(), B.stride(0),
                              None if offset is None else offset.data_ptr(), Y.data_ptr(), n,
                              P, n, K, code, st)
    N.check(rc, "contract")
    return Y


def _ref(A, B, offset):
    y = A.double().t() @ B.double()
    if offset is not None:
        y = y + offset
    scale = A.double().abs().t() @ B.double().abs()
    return y, scale


SHAPES = [(1, 1, 1), (5, 3, 15), (64, 16, 16), (100, 33, 17), (257, 64, 300), (1000, 100, 1000),
          (130, 129, 4096), (2000, 1024, 333), (64, 7, 0)]


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-13), (torch.float32, 2e-6)])
@pytest.mark.parametrize("P,n,K", SHAPES)
def test_contract_matches_fp64_product(gpu, dtype, tol, P, n, K):
    g = torch.Generator(device="cuda").manual_seed(P * 7919 + n * 31 + K)
    A = torch.randn((K, P), dtype=torch.float64, device="cuda", generator=g).to(dtype)
    B = torch.randn((K, n), dtype=torch.float64, device="cuda", generator=g).to(dtype)
    off = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    Y = contract(A, B, off)
    ref, scale = _ref(A, B, off)
    assert torch.isfinite(Y).all()
    err = ((Y - ref).abs() / (scale + 1e-300)).max().item() if K else (Y - ref).abs().max().item()
    assert err <= tol, (P, n, K, err)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_contract_strided_operands(gpu, dtype):
    """Row strides larger than the logical widths (lda > P, ldb > n) and no offset."""
    g = torch.Generator(device="cuda").manual_seed(5)
    A_full = torch.randn((700, 300), dtype=torch.float64, device="cuda", generator=g).to(dtype)
    B_full = torch.randn((700, 90), dtype=torch.float64, device="cuda", generator=g).to(dtype)
    A, B = A_full[:, :257], B_full[:, :70]
    Y = contract(A, B, None, P=257, n=70)
    ref, scale = _ref(A, B, None)
    assert ((Y - ref).abs() / scale).max().item() <= (1e-13 if dtype == torch.float64 else 2e-6)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_contract_row_independent_of_batch(gpu, dtype):
    """A point's row is the same bit for bit alone, at another batch position and in a batch."""
    g = torch.Generator(device="cuda").manual_seed(11)
    K, P, n = 1500, 3000, 200
    A = torch.randn((K, P), dtype=torch.float64, device="cuda", generator=g).to(dtype)
    B = torch.randn((K, n), dtype=torch.float64, device="cuda", generator=g).to(dtype)
    off = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    Y = contract(A, B, off)
    for lo, hi in [(0, 1), (17, 18), (100, 163), (2999, 3000), (5, 1000)]:
        Ys = contract(A[:, lo:hi].contiguous(), B, off)
        assert torch.equal(Ys, Y[lo:hi]), (lo, hi)


def test_contract_rejects_bad_arguments(gpu):
    A = torch.zeros((4, 8), dtype=torch.float32, device="cuda")
    B = torch.zeros((4, 8), dtype=torch.float32, device="cuda")
    Y = torch.zeros((8, 8), dtype=torch.float64, device="cuda")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    f = N.fn["sbi_contract"]
    assert f(A.data_ptr(), 4, B.data_ptr(), 8, None, Y.data_ptr(), 8, 8, 8, 4, 1, st) == N.SBI_ERR_INVALID
    assert f(A.data_ptr(), 8, B.data_ptr(), 8, None, Y.data_ptr(), 8, 8, 8, 4, 5, st) == N.SBI_ERR_INVALID
    assert f(A.data_ptr(), 8, B.data_ptr(), 8, None, Y.data_ptr(), 8, 8, 8, 4, 1, st) == N.SBI_OK

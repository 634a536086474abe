"""GPU: sbi_contract (Y = offset + A^T B on the matrix cores) through the C-ABI.

Reference: the same product in FP64 on the device with torch (operands upcast exactly).  Bounds,
per element, relative to sum_k |A[k,p]| |B[k,j]|: FP64 operands 1e-13 (FP64 MFMA, K <= 4096 terms);
FP32 operands 2e-6 (products and 256-term partial sums in FP32, flushed into FP64).  Rows are
independent: a point's row is identical bit for bit alone and inside a larger batch.
sbi_contract_split_k4(_ws): the same bits as the plain layout and the unsplit kernel.
sbi_contract_bf16_k8: against the FP64 product of the same BF16 operands (bound from the FP32
partial sums of 1024 terms), and the same bits for its three tile shapes.
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


def contract_split(A64, B64, A32, B32, offset=None):
    K64, P = A64.shape
    K32 = A32.shape[0]
    n = B64.shape[1]
    Y = torch.full((P, n), float("nan"), dtype=torch.float64, device=A64.device)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = N.fn["sbi_contract_split"](A64.data_ptr(), A64.stride(0), B64.data_ptr(), B64.stride(0),
                                    K64, A32.data_ptr() if K32 else None, P,
                                    B32.data_ptr() if K32 else None, n, K32,
                                    None if offset is None else offset.data_ptr(), Y.data_ptr(), n,
                                    P, n, st)
    N.check(rc, "contract_split")
    return Y


@pytest.mark.parametrize("P,n,K64,K32", [(1, 1, 1, 0), (70, 33, 64, 0), (130, 100, 64, 500),
                                         (257, 129, 13, 1000), (64, 16, 64, 17)])
def test_contract_split_matches_fp64_product(gpu, P, n, K64, K32):
    """FP64 head rows on the FP64 MFMA (bound 1e-13 when there is no FP32 part: this also checks
    the lane permutation that maps the FP64 result into the FP32 layout), FP32 rows as sbi_contract
    (bound 2e-6 of sum |a||b|)."""
    g = torch.Generator(device="cuda").manual_seed(P + 7 * n + 31 * K64 + K32)
    A64 = torch.randn((K64, P), dtype=torch.float64, device="cuda", generator=g)
    B64 = torch.randn((K64, n), dtype=torch.float64, device="cuda", generator=g)
    A32 = torch.randn((max(K32, 1), P), dtype=torch.float64, device="cuda", generator=g).float()[:K32]
    B32 = torch.randn((max(K32, 1), n), dtype=torch.float64, device="cuda", generator=g).float()[:K32]
    off = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    Y = contract_split(A64, B64, A32.contiguous(), B32.contiguous(), off)
    ref = off + A64.t() @ B64 + A32.double().t() @ B32.double()
    s64 = A64.abs().t() @ B64.abs()
    s32 = A32.double().abs().t() @ B32.double().abs()
    assert torch.isfinite(Y).all()
    assert ((Y - ref).abs() <= 1e-13 * s64 + 2e-6 * s32 + 1e-300).all()


def test_contract_split_rows_independent_of_batch(gpu):
    g = torch.Generator(device="cuda").manual_seed(3)
    K64, K32, P, n = 64, 700, 500, 90
    A64 = torch.randn((K64, P), dtype=torch.float64, device="cuda", generator=g)
    B64 = torch.randn((K64, n), dtype=torch.float64, device="cuda", generator=g)
    A32 = torch.randn((K32, P), dtype=torch.float32, device="cuda", generator=g)
    B32 = torch.randn((K32, n), dtype=torch.float32, device="cuda", generator=g)
    Y = contract_split(A64, B64, A32, B32)
    for lo, hi in [(0, 1), (3, 4), (100, 211), (499, 500)]:
        Ys = contract_split(A64[:, lo:hi].contiguous(), B64, A32[:, lo:hi].contiguous(), B32)
        assert torch.equal(Ys, Y[lo:hi]), (lo, hi)


def _k4(M, rows):
    """(rows, m) plain layout -> k-interleaved (rows / 4, m, 4), rows a multiple of 16."""
    return M.view(rows // 4, 4, M.shape[1]).permute(0, 2, 1).contiguous()


@pytest.mark.parametrize("P,n,K64,K32", [(64, 32, 0, 16), (100, 70, 5, 48), (333, 257, 64, 1024),
                                         (17, 3, 3, 0)])
def test_contract_split_k4_equals_plain_layout(gpu, P, n, K64, K32):
    """The k-interleaved operands give the same sums in the same order: bit-identical to
    sbi_contract_split on the same values (K32 a multiple of 16, ragged P and n)."""
    from smolyax_amd import _native as N

    g = torch.Generator(device="cuda").manual_seed(P + n + K32)
    A64 = torch.randn((max(K64, 1), P), dtype=torch.float64, device="cuda", generator=g)[:K64]
    B64 = torch.randn((max(K64, 1), n), dtype=torch.float64, device="cuda", generator=g)[:K64]
    A32 = torch.randn((max(K32, 16), P), dtype=torch.float32, device="cuda", generator=g)[:K32]
    B32 = torch.randn((max(K32, 16), n), dtype=torch.float32, device="cuda", generator=g)[:K32]
    off = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    Y0 = contract_split(A64.contiguous(), B64.contiguous(), A32.contiguous(), B32.contiguous(), off)
    A4 = _k4(A32.contiguous(), K32) if K32 else A32
    B4 = _k4(B32.contiguous(), K32) if K32 else B32
    Y1 = torch.empty((P, n), dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    N.check(N.fn["sbi_contract_split_k4"](
        A64.data_ptr() if K64 else None, P, B64.data_ptr() if K64 else None, n, K64,
        A4.data_ptr() if K32 else None, P, B4.data_ptr() if K32 else None, n, K32, off.data_ptr(),
        Y1.data_ptr(), n, P, n, st), "contract_k4")
    torch.cuda.synchronize()
    assert torch.equal(Y0, Y1)


@pytest.mark.parametrize("P,n,K64,K32,split", [(64, 761, 256, 3792, True), (256, 43, 0, 1264, True),
                                               (200, 1, 64, 512, True), (500, 1024, 256, 2048, True),
                                               (33, 100, 7, 16, True), (5000, 1024, 64, 1024, False),
                                               (64, 64, 0, 256, False), (7, 5, 300, 0, False)])
def test_contract_split_k4_ws_same_bits(gpu, P, n, K64, K32, split):
    """Small batches split k over extra workgroups (256-k partial tiles and the FP64 head in the
    workspace, added in the unsplit order by a second kernel): the same bits as
    sbi_contract_split_k4, for ragged P and n, K a multiple of 256 or not, with and without a
    head; shapes that fill the GPU (or have one chunk) need no workspace and run unsplit."""
    print(f"split_k4_ws P={P} n={n} K64={K64} K32={K32}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(P + 3 * n + K32)
    A64 = torch.randn((max(K64, 1), P), dtype=torch.float64, device="cuda", generator=g)[:K64].contiguous()
    B64 = torch.randn((max(K64, 1), n), dtype=torch.float64, device="cuda", generator=g)[:K64].contiguous()
    A32 = torch.randn((max(K32, 16), P), dtype=torch.float32, device="cuda", generator=g)[:K32]
    B32 = torch.randn((max(K32, 16), n), dtype=torch.float32, device="cuda", generator=g)[:K32]
    off = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    A4 = _k4(A32.contiguous(), K32) if K32 else A32
    B4 = _k4(B32.contiguous(), K32) if K32 else B32
    args = (A64.data_ptr() if K64 else None, P, B64.data_ptr() if K64 else None, n, K64,
            A4.data_ptr() if K32 else None, P, B4.data_ptr() if K32 else None, n, K32, off.data_ptr())
    st = torch.cuda.current_stream().cuda_stream
    Y0 = torch.full((P, n), float("nan"), dtype=torch.float64, device="cuda")
    N.check(N.fn["sbi_contract_split_k4"](*args, Y0.data_ptr(), n, P, n, st), "k4")
    ws = N.fn["sbi_contract_split_k4_workspace"](P, n, K64, K32)
    assert (ws > 0) == split, ws
    W = torch.empty(max(ws, 1), dtype=torch.uint8, device="cuda")
    Y1 = torch.full((P, n), float("nan"), dtype=torch.float64, device="cuda")
    N.check(N.fn["sbi_contract_split_k4_ws"](*args, Y1.data_ptr(), n, P, n, W.data_ptr(), ws, st),
            "k4_ws")
    # a workspace too small runs the unsplit kernel
    Y2 = torch.full((P, n), float("nan"), dtype=torch.float64, device="cuda")
    N.check(N.fn["sbi_contract_split_k4_ws"](*args, Y2.data_ptr(), n, P, n, W.data_ptr(),
                                             max(ws - 256, 0), st), "k4_ws_small")
    torch.cuda.synchronize()
    assert torch.equal(Y0, Y1)
    assert torch.equal(Y0, Y2)


def _k8(Mb, K):
    """(K, m) BF16 -> sbi_contract_bf16_k8's k-interleaved layout [K/8][m][8]."""
    return Mb.view(K // 8, 8, Mb.shape[1]).permute(0, 2, 1).contiguous()


def _bf16_contract(A64, B64, K64, A8, B8, K16, off, P, n, lda, ldb):
    Y = torch.full((P, n), float("nan"), dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    N.check(N.fn["sbi_contract_bf16_k8"](A64.data_ptr() if K64 else None, lda,
                                         B64.data_ptr() if K64 else None, ldb, K64,
                                         A8.data_ptr() if K16 else None, lda,
                                         B8.data_ptr() if K16 else None, ldb, K16,
                                         off.data_ptr(), Y.data_ptr(), n, P, n, st), "bf16_k8")
    torch.cuda.synchronize()
    return Y


@pytest.mark.parametrize("P,n,K64,K16", [(100, 33, 0, 64), (257, 300, 5, 1024), (1000, 130, 16, 2080),
                                         (4096, 4096, 3, 1024), (4000, 4100, 2, 96), (64, 7, 9, 0),
                                         (5, 3, 0, 32)])
def test_contract_bf16_matches_fp64_product(gpu, P, n, K64, K16):
    """sbi_contract_bf16_k8 against the FP64 product of the same BF16 operands (and FP64 head
    rows); 4000 x 4100 takes the 4 x 4 tiles with partial tiles on both edges.  BF16 products are exact in FP32; each FP32 partial sum covers at most 1024 terms
    before it is added into FP64, so per element |Y - ref| <= 1023 * 2^-24 * sum |a b| (+ the
    FP64 head's rounding): bound 6.2e-5 relative to sum_k |A[k,p]| |B[k,j]|."""
    print(f"bf16_k8 P={P} n={n} K64={K64} K16={K16}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(P * 13 + n + K16)
    A64 = torch.randn((max(K64, 1), P), dtype=torch.float64, device="cuda", generator=g)[:K64].contiguous()
    B64 = torch.randn((max(K64, 1), n), dtype=torch.float64, device="cuda", generator=g)[:K64].contiguous()
    Ab = torch.randn((max(K16, 8), P), dtype=torch.float64, device="cuda", generator=g)[:K16].to(torch.bfloat16)
    Bb = torch.randn((max(K16, 8), n), dtype=torch.float64, device="cuda", generator=g)[:K16].to(torch.bfloat16)
    off = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    A8 = _k8(Ab, K16) if K16 else Ab
    B8 = _k8(Bb, K16) if K16 else Bb
    Y = _bf16_contract(A64, B64, K64, A8, B8, K16, off, P, n, P, n)
    Ad = torch.cat([A64, Ab.double()]) if K16 else A64
    Bd = torch.cat([B64, Bb.double()]) if K16 else B64
    ref = Ad.t() @ Bd + off
    scale = Ad.abs().t() @ Bd.abs()
    assert torch.isfinite(Y).all()
    err = ((Y - ref).abs() / (scale + 1e-300)).max().item()
    assert err <= 6.2e-5, (P, n, K64, K16, err)


def test_contract_bf16_tile_shapes_same_bits(gpu):
    """The three tile shapes of sbi_contract_bf16_k8 (4 x 4 fragments per wave for grids of at
    least 1024 workgroups of 128 x 128, else 4 x 2 for 1024 of 128 x 64, else 2 x 2) sum every
    element in the same order: 4096, 2048 and 512 points of the same operands (one row stride)
    give the same bits on their common rows, with and without FP64 head rows."""
    n, K16 = 4096, 2080
    g = torch.Generator(device="cuda").manual_seed(11)
    Ab = torch.randn((K16, 4096), dtype=torch.float64, device="cuda", generator=g).to(torch.bfloat16)
    Bb = torch.randn((K16, n), dtype=torch.float64, device="cuda", generator=g).to(torch.bfloat16)
    A8, B8 = _k8(Ab, K16), _k8(Bb, K16)
    off = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    for K64 in (0, 7):
        A64 = torch.randn((max(K64, 1), 4096), dtype=torch.float64, device="cuda", generator=g)[:K64].contiguous()
        B64 = torch.randn((max(K64, 1), n), dtype=torch.float64, device="cuda", generator=g)[:K64].contiguous()
        Ys = [_bf16_contract(A64, B64, K64, A8, B8, K16, off, P, n, 4096, n) for P in (4096, 2048, 512)]
        assert torch.equal(Ys[0][:2048], Ys[1]), K64
        assert torch.equal(Ys[0][:512], Ys[2]), K64

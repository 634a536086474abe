"""GPU: set-up on the device through the C-ABI.

sbi_pack_rows_f64 (F's rows from the f values at the distinct grid nodes, zeta folded in) and
sbi_hierarchize_f64 (grid values -> Newton surpluses) against their host forms
(packing.pack_rows, newton.apply_steps / newton.hierarchize), bit for bit: both kernels round
every product and difference on its own, in the host order.  Then set_f end to end: the device F,
the record program's F stream and the Newton surpluses equal the host packing's, for nested
(Leja) and non-nested (Gauss-Hermite, proxy Newton form) nodes.
"""

import ctypes

import numpy as np
import pytest
import torch

from oracle import indices as OI
from smolyax_amd import _native as N

pytestmark = pytest.mark.gpu


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _pack(V, node, scale, n_rows=None, out=None):
    d_out = V.shape[1]
    n_rows = len(node) if n_rows is None else n_rows
    F = torch.full((max(n_rows, 1), d_out), float("nan"), dtype=torch.float64, device=V.device) \
        if out is None else out
    rc = N.fn["sbi_pack_rows_f64"](V.data_ptr(), V.shape[0], d_out, node.data_ptr(),
                                   scale.data_ptr(), n_rows, F.data_ptr(), _stream())
    N.check(rc, "pack_rows")
    torch.cuda.synchronize()
    return F


@pytest.mark.parametrize("n_nodes,d_out,n_rows", [(50, 1, 333), (17, 3, 100), (40, 8, 1000),
                                                  (9, 257, 64), (30, 1024, 200), (1, 2, 1)])
def test_pack_rows_matches_host(gpu, n_nodes, d_out, n_rows):
    from smolyax_amd import packing

    print(f"pack_rows n_nodes={n_nodes} d_out={d_out} n_rows={n_rows}", flush=True)
    rng = np.random.default_rng(n_nodes * 1000 + d_out)
    V = rng.standard_normal((n_nodes, d_out)) * 10.0 ** rng.integers(-5, 5, (n_nodes, 1))
    node = rng.integers(0, n_nodes, n_rows).astype(np.int32)
    node[::7] = -1                       # padding rows
    node[3::11] = n_nodes                # one past the table: also a zero row
    scale = rng.integers(-3000, 3000, n_rows).astype(np.float64)
    ref = packing.pack_rows(V, node, scale)
    Vd = torch.as_tensor(V, device=gpu)
    F = _pack(Vd, torch.as_tensor(node, device=gpu), torch.as_tensor(scale, device=gpu))
    assert torch.equal(F[:n_rows].cpu(), torch.as_tensor(ref))


def test_pack_rows_unaligned_and_empty(gpu):
    """An output at an odd element offset (8-byte loads and stores) and zero rows."""
    from smolyax_amd import packing

    rng = np.random.default_rng(7)
    V = rng.standard_normal((20, 6))
    node = rng.integers(-1, 20, 50).astype(np.int32)
    scale = rng.standard_normal(50)
    buf = torch.full((50 * 6 + 1,), float("nan"), dtype=torch.float64, device=gpu)
    out = buf[1:].view(50, 6)
    _pack(torch.as_tensor(V, device=gpu), torch.as_tensor(node, device=gpu),
          torch.as_tensor(scale, device=gpu), out=out)
    assert torch.equal(out.cpu(), torch.as_tensor(packing.pack_rows(V, node, scale)))
    assert torch.isnan(buf[0]).item()
    # no rows: nothing written
    F = torch.full((1, 6), 5.0, dtype=torch.float64, device=gpu)
    nd = torch.zeros(1, dtype=torch.int32, device=gpu)
    sc = torch.ones(1, dtype=torch.float64, device=gpu)
    _pack(torch.as_tensor(V, device=gpu), nd, sc, n_rows=0, out=F)
    assert torch.equal(F.cpu(), torch.full((1, 6), 5.0, dtype=torch.float64))


def _leja_nus(d, m):
    from smolyax_amd import indices

    k = np.log((2 + np.arange(d)) / np.log(2))
    t = OI.find_approximate_threshold(k, m, True)
    return indices.indexset(k, t)


@pytest.mark.parametrize("d,m,d_out", [(3, 60, 1), (6, 300, 5), (12, 800, 300), (40, 2000, 2)])
def test_hierarchize_matches_host(gpu, d, m, d_out):
    from smolyax_amd import newton
    from smolyax_amd import nodes as SN

    print(f"hierarchize d={d} m={m} d_out={d_out}", flush=True)
    nus = _leja_nus(d, m)
    gen = SN.Leja(dim=d)
    steps = newton.hierarchy_steps(nus, gen)
    rng = np.random.default_rng(d + d_out)
    vals = rng.standard_normal((len(nus), d_out))
    ref = newton.hierarchize(nus, vals, gen)
    S = torch.as_tensor(vals, device=gpu).contiguous()
    t = [torch.as_tensor(a, device=gpu) for a in (steps.step_off, steps.tgt, steps.src_off,
                                                  steps.src, steps.coef)]
    rc = N.fn["sbi_hierarchize_f64"](S.data_ptr(), S.shape[0], d_out, d_out, t[0].data_ptr(),
                                     steps.n_steps, t[1].data_ptr(), t[2].data_ptr(),
                                     t[3].data_ptr(), t[4].data_ptr(), _stream())
    N.check(rc, "hierarchize")
    torch.cuda.synchronize()
    assert torch.equal(S.cpu(), torch.as_tensor(ref))


def test_hierarchize_wider_rows(gpu):
    """ld > d_out: the columns beyond d_out are left alone."""
    from smolyax_amd import newton
    from smolyax_amd import nodes as SN

    nus = _leja_nus(5, 120)
    gen = SN.Leja(dim=5)
    steps = newton.hierarchy_steps(nus, gen)
    vals = np.random.default_rng(3).standard_normal((len(nus), 4))
    S = torch.as_tensor(vals, device=gpu).contiguous()
    t = [torch.as_tensor(a, device=gpu) for a in (steps.step_off, steps.tgt, steps.src_off,
                                                  steps.src, steps.coef)]
    N.check(N.fn["sbi_hierarchize_f64"](S.data_ptr(), S.shape[0], 3, 4, t[0].data_ptr(),
                                        steps.n_steps, t[1].data_ptr(), t[2].data_ptr(),
                                        t[3].data_ptr(), t[4].data_ptr(), _stream()), "hier")
    torch.cuda.synchronize()
    ref = newton.hierarchize(nus, vals[:, :3], gen)
    assert torch.equal(S[:, :3].cpu(), torch.as_tensor(ref))
    assert torch.equal(S[:, 3].cpu(), torch.as_tensor(vals[:, 3]))


@pytest.mark.parametrize("kind,d,d_out,m", [("leja", 6, 1, 300), ("leja", 8, 7, 500),
                                             ("gh", 4, 1, 200), ("gh", 5, 3, 150)])
def test_set_f_device_layout_matches_host(gpu, kind, d, d_out, m):
    from smolyax_amd.interpolation import SmolyakBarycentricInterpolator
    from smolyax_amd import newton
    from smolyax_amd import nodes as SN
    from smolyax_amd import packing

    print(f"set_f {kind} d={d} d_out={d_out} m={m}", flush=True)
    k = np.log((2 + np.arange(d)) / np.log(2))
    gen = SN.Leja(dim=d) if kind == "leja" else SN.GaussHermite(dim=d)
    t = OI.find_approximate_threshold(k, m, gen.is_nested)
    A = np.random.default_rng(d).standard_normal((d_out, d))
    f = lambda x: np.sin(A @ x + 0.3)
    ip = SmolyakBarycentricInterpolator(d_out=d_out, node_gen=gen, k=k, t=t)
    fe = ip.set_f(f=f)
    blocks = packing.nonzero_blocks(k, t)
    L = packing.build_layout(blocks, gen, d, d_out)
    F, off, _ = packing.assemble_values(L, blocks, gen, f, {})
    assert torch.equal(ip._dev["F"].cpu(), torch.as_tensor(F))
    assert torch.equal(ip._dev["offset"].cpu(), torch.as_tensor(off))
    if d_out == 1:
        assert torch.equal(ip._dev["pF"].cpu(), torch.as_tensor(packing.program_values(ip._prog, F)))
    if kind == "leja":
        nf = ip._nform
        vals = packing.stack_values([fe[nu] for nu in nf.nus], d_out)
        assert np.array_equal(nf.s, newton.hierarchize(nf.nus, vals, nf.node_gen))
    else:
        # proxy Newton form: the host program on the same proxy-grid values gives the same bits
        nf = ip._nform
        assert nf is not None
        zero = np.array([g(0)[0] for g in nf.node_gen], dtype=np.float64)
        X = np.stack([nf.grid_point(nu, zero) for nu in nf.nus])
        Y = ip.evaluate(torch.as_tensor(X, device=gpu), precision="fp64").cpu().numpy()
        assert np.array_equal(nf.s, newton.hierarchize(nf.nus, Y, nf.node_gen))

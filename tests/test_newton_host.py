"""Hierarchical (Newton) form on the host, no GPU: surpluses and the device layout reproduce the
oracle; the kernel's mixed-precision arithmetic (modelled in numpy) stays within 1e-6."""

import numpy as np
import pytest

from oracle import indices as OI
from oracle import nodes as ON
from oracle import smolyak as OS


@pytest.mark.parametrize("d,m,dom", [(1, 9, False), (3, 40, True), (6, 200, False), (10, 1000, False)])
def test_newton_layout_matches_oracle(d, m, dom):
    from smolyax_amd import newton as NW
    from smolyax_amd import nodes as SN

    rng = np.random.default_rng(d)
    k = np.log((2 + np.arange(d)) / np.log(2))
    if dom:
        dm = np.sort(rng.random((d, 2)), axis=1) + [0, 0.3]
        ong, sng = ON.leja(domains=dm), SN.Leja(domains=dm)
    else:
        ong, sng = ON.leja(dim=d), SN.Leja(dim=d)
    t = OI.find_approximate_threshold(k, m, True)
    a = rng.standard_normal(d) / (np.arange(d) + 1.0)
    f = lambda x: np.exp(-(np.asarray(x) @ a))
    ref = OS.Interpolant(d_out=1, node_gen=ong, k=k, t=t, f=f)
    nf = NW.NewtonForm(k, t, sng, 1)
    zero = np.array([g(0)[0] for g in sng])
    X = np.stack([nf.grid_point(nu, zero) for nu in nf.nus])
    nf.set_values(f(X).reshape(-1, 1))
    L = NW.build_newton_layout(nf)
    x = np.stack([g.scale(c) for g, c in zip(ong, rng.uniform(-1, 1, (64, d)).T)], axis=1)
    yr = ref(x)[:, 0]
    y64 = NW.eval_newton_layout(L, x, fp32=False)
    y32 = NW.eval_newton_layout(L, x, fp32=True)
    scale = np.max(np.abs(yr))
    assert np.max(np.abs(y64 - yr)) / scale < 1e-12
    assert np.max(np.abs(y32 - yr)) / scale < 1e-6
    # grid nodes are interpolated
    assert np.max(np.abs(NW.eval_newton_layout(L, X[:50]) - f(X[:50]))) < 1e-12 * max(1, np.max(np.abs(f(X))))
    # every node is in exactly one group and its parent row precedes it in depth
    g4 = L.grp.reshape(-1, 4)
    assert g4[0, 0] == 0 and g4[-1, 1] == L.n_nodes and np.all(g4[1:, 0] == g4[:-1, 1])


def test_hierarchize_polynomial_reproduction():
    """Surpluses of a polynomial in P_Lambda vanish outside its support."""
    from smolyax_amd import newton as NW
    from smolyax_amd import nodes as SN

    d = 3
    k = np.log((2 + np.arange(d)) / np.log(2))
    t = OI.find_approximate_threshold(k, 60, True)
    nf = NW.NewtonForm(k, t, SN.Leja(dim=d), 1)
    X = np.stack([nf.grid_point(nu, np.zeros(d)) for nu in nf.nus])
    f = lambda x: 1.0 + 2.0 * x[:, 0] - 0.5 * x[:, 1] * x[:, 0]
    nf.set_values(f(X).reshape(-1, 1))
    nz = {nu for nu, s in zip(nf.nus, nf.s[:, 0]) if abs(s) > 1e-13}
    assert nz <= {(), ((0, 1),), ((0, 1), (1, 1))}


def test_newton_layout_limits_raise_capacity_error():
    """A node with more than 8 factors cannot be described for the basis kernel: the host raises
    CapacityError (set_f then keeps the barycentric paths only)."""
    from types import SimpleNamespace

    from smolyax_amd import newton

    d = 10
    nus = [(), tuple((j, 1) for j in range(9))]
    nf = SimpleNamespace(d_in=d, maxdeg=np.array([1] * 9 + [0]), nus=nus)
    L = SimpleNamespace()
    with pytest.raises(newton.CapacityError):
        newton.build_basis_tables(nf, L)
    assert issubclass(newton.CapacityError, ValueError)


def test_gemm_row_order_puts_heavy_rows_first():
    """The GEMM paths' basis rows are a permutation of nf.nus with the head rows (largest
    max |h| * ||s||) first; basis descriptors and head factor tables follow that order."""
    from smolyax_amd import newton
    from smolyax_amd import nodes as SN
    from smolyax_amd import indices as SI

    d = 12
    k = np.log((2 + np.arange(d)) / np.log(2))
    t = SI.find_approximate_threshold(k, 400, True)
    nf = newton.NewtonForm(k, t, SN.Leja(dim=d), 3)
    rng = np.random.default_rng(0)
    nf.s = rng.standard_normal((len(nf.nus), 3)) / (1.0 + np.arange(len(nf.nus)))[:, None]
    L = newton.build_newton_layout(nf)
    order = L.basis_order
    assert sorted(order.tolist()) == list(range(len(nf.nus)))
    assert L.n_head == min(newton.HEAD_ROWS, len(nf.nus))
    # descriptor row r decodes to the factors of nus[order[r]]
    desc = L.basis_desc.view(np.uint16).reshape(L.basis_n, 8)
    tab_of = {}
    for i, j in enumerate(L.tab_dim[: L.n_tab_dims]):
        for kk in range(1, int(L.tab_kmax[i]) + 1):
            tab_of[int(L.tab_row[i]) + kk - 1] = (int(j), kk)
    for r in range(L.basis_n):
        got = sorted(tab_of[int(v)] for v in desc[r] if v != 0xFFFF)
        assert got == sorted(nf.nus[order[r]])
    fac = L.head_fac.reshape(-1, 8, 2)
    for r in range(L.n_head):
        got = sorted((int(j), int(kk)) for j, kk in fac[r] if j >= 0)
        assert got == sorted(nf.nus[order[r]])


@pytest.mark.parametrize("d,m", [(1, 9), (3, 40), (10, 1000), (200, 5000), (1000, 10000)])
def test_matrix_core_layout_matches_oracle(d, m):
    """Tables of the matrix-core mixed kernel: cost-ordered rows of W, dense surplus vectors in
    tiles of 16 groups (A steps in MFMA lane order), work items and their wave groups.  The numpy
    model of its arithmetic matches the oracle (FP64 Newton layout at large d) to 1e-6; every
    node's folded surplus appears exactly once in the A steps; every tile's steps are covered by
    its items exactly once; the wave groups cover the items in order."""
    from smolyax_amd import newton as NW
    from smolyax_amd import nodes as SN

    rng = np.random.default_rng(d)
    k = np.log((2 + np.arange(d)) / np.log(2))
    t = OI.find_approximate_threshold(k, m, True)
    a = rng.standard_normal(d) / (np.arange(d) + 1.0) ** 2
    f = lambda x: np.exp(-(np.asarray(x) @ a))
    sng = SN.Leja(dim=d)
    nf = NW.NewtonForm(k, t, sng, 1)
    X = np.stack([nf.grid_point(nu, np.zeros(d)) for nu in nf.nus])
    nf.set_values(f(X).reshape(-1, 1))
    L = NW.build_newton_layout(nf)
    assert L.mt_a is not None
    x = rng.uniform(-1, 1, (32, d))
    y = NW.eval_mfma_layout(L, x)
    if d <= 10:
        ref = OS.Interpolant(d_out=1, node_gen=ON.leja(dim=d), k=k, t=t, f=f)
        yr = ref(x)[:, 0]
    else:
        yr = NW.eval_newton_layout(L, x, fp32=False)
    assert np.max(np.abs(y - yr)) / np.max(np.abs(yr)) < 1e-6
    # each node's folded surplus once: the nonzero A entries are the nodes' c_k * s (nonzero ones)
    g4 = L.grp.reshape(-1, 4)
    cnt = np.diff(g4[:, :2], axis=1)[:, 0]
    folded = L.node_s32.astype(np.float64) * L.dim_c[g4[:, 2].repeat(cnt), g4[:, 3].repeat(cnt)]
    assert np.count_nonzero(L.mt_a) == np.count_nonzero(folded)
    # items cover each tile's steps once, in A-step order
    rec = L.mi_rec.reshape(-1, 4)
    cover = np.zeros(L.mt_steps, dtype=int)
    for a0, w0, ns, tt in rec[: L.mi_n]:
        cover[a0: a0 + ns] += 1
        assert 0 <= tt < L.mt_n and w0 >= 0
    assert np.all(cover == 1)
    assert L.mwave_item[0] == 0 and L.mwave_item[-1] == L.mi_n
    assert np.all(np.diff(L.mwave_item) >= 0) and len(L.mwave_item) == NW.MFMA_WAVES + 1
    # the depth levels of the cost-ordered records are those of level_off
    depth = [len(v) for v in sorted({nu[:-1] for nu in nf.nus if len(nu) > 1}, key=len)]
    assert list(np.diff(L.level_off)) == [depth.count(dd) for dd in sorted(set(depth))]


@pytest.mark.parametrize("d,m", [(1, 9), (4, 120), (10, 1000), (60, 3000)])
def test_hierarchy_steps_program(d, m):
    """The step program that sbi_hierarchize_f64 runs: every multi-index is a target once per
    active dim, in ascending dim then level order; a target's sources are the multi-indices with
    its level along the step's dim lowered to 0 .. lv - 1 (in that order); no step reads a row it
    writes (the kernel's precondition); applied to the grid values it gives surpluses whose
    Newton form interpolates them (checked at every grid node)."""
    from smolyax_amd import indices as SI
    from smolyax_amd import newton as NW
    from smolyax_amd import nodes as SN

    k = np.log((2 + np.arange(d)) / np.log(2))
    t = OI.find_approximate_threshold(k, m, True)
    nus = SI.indexset(k, t)
    gen = SN.Leja(dim=d)
    st = NW.hierarchy_steps(nus, gen)
    assert st.step_off[0] == 0 and st.step_off[-1] == len(st.tgt) and len(st.src_off) == len(st.tgt) + 1
    assert len(st.tgt) == sum(len(nu) for nu in nus)
    assert st.src_off[-1] == len(st.src) == len(st.coef) == sum(v for nu in nus for _, v in nu)
    prev = (-1, 0)
    for s in range(st.n_steps):
        tg = st.tgt[st.step_off[s]:st.step_off[s + 1]]
        assert len(tg) > 0 and len(set(tg.tolist())) == len(tg)
        srcs = set(st.src[st.src_off[st.step_off[s]]:st.src_off[st.step_off[s + 1]]].tolist())
        assert not (srcs & set(tg.tolist()))
        # one (dim, level) per step, in ascending order: the dim that the first source drops
        i0 = st.step_off[s]
        nu0, low0 = dict(nus[st.tgt[i0]]), dict(nus[st.src[st.src_off[i0]]])
        dropped = set(nu0) - set(low0)
        assert len(dropped) == 1
        jd = dropped.pop()
        key = (jd, nu0[jd])
        assert key > prev
        prev = key
        for i in range(st.step_off[s], st.step_off[s + 1]):
            nu = dict(nus[st.tgt[i]])
            lv = nu[key[0]]
            assert lv == key[1]
            q0, q1 = st.src_off[i], st.src_off[i + 1]
            assert q1 - q0 == lv
            for l, q in enumerate(range(q0, q1)):
                low = dict(nu)
                if l == 0:
                    low.pop(key[0])
                else:
                    low[key[0]] = l
                assert nus[st.src[q]] == tuple(sorted(low.items()))
    # interpolation at the grid nodes
    rng = np.random.default_rng(d)
    vals = rng.standard_normal((len(nus), 2))
    s = NW.apply_steps(st, vals)
    nf = NW.NewtonForm(k, t, gen, 2)
    nf.s = s
    zero = np.array([g(0)[0] for g in gen])
    X = np.stack([nf.grid_point(nu, zero) for nu in nus])
    assert np.max(np.abs(nf.evaluate(X) - vals)) < 1e-9 * max(1.0, np.max(np.abs(vals)))

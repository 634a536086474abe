"""Host packing (no GPU): the ragged device layout evaluated in numpy equals the oracle."""

import numpy as np
import pytest

import layout_eval
from oracle import indices as OI
from oracle import nodes as ON
from oracle import smolyak as OS


@pytest.mark.parametrize("kind,d,d_out,m", [("leja", 6, 3, 200), ("gh", 5, 2, 300), ("leja", 2, 1, 40),
                                             ("leja_dom", 4, 2, 90), ("gh_ms", 3, 2, 120)])
def test_layout_matches_oracle(kind, d, d_out, m):
    from smolyax_amd import nodes as SN
    from smolyax_amd import packing

    rng = np.random.default_rng(d)
    k = np.log((2 + np.arange(d)) / np.log(2))
    if kind == "leja":
        ong, sng = ON.leja(dim=d), SN.Leja(dim=d)
    elif kind == "leja_dom":
        dom = np.sort(rng.random((d, 2)), axis=1) + [0, 0.2]
        ong, sng = ON.leja(domains=dom), SN.Leja(domains=dom)
    elif kind == "gh":
        ong, sng = ON.gauss_hermite(dim=d), SN.GaussHermite(dim=d)
    else:
        mean, sc = rng.standard_normal(d), rng.random(d) + 0.5
        ong, sng = ON.gauss_hermite(mean=mean, scaling=sc), SN.GaussHermite(mean, sc)
    t = OI.find_approximate_threshold(k, m, ong.is_nested)
    A = rng.standard_normal((d_out, d))
    f = lambda x: np.sin(A @ x + 0.3)
    ref = OS.Interpolant(d_out=d_out, node_gen=ong, k=k, t=t, f=f)
    blocks = packing.nonzero_blocks(k, t)
    L = packing.build_layout(blocks, sng, d, d_out)
    F, off, n_new = packing.assemble_values(L, blocks, sng, f, {})
    assert n_new == ref.n_f_evals_new == ref.n_f_evals
    assert L.k_raw == sum(int(np.prod([v + 1 for v in g])) for g in L.blk_degs)
    x = np.stack([g.scale(c) for g, c in zip(ong, rng.uniform(-0.9, 0.9, (40, d)).T)], axis=1)
    y = layout_eval.eval_layout(L, F, off, x)
    yr = ref(x)
    assert np.max(np.abs(y - yr)) / np.max(np.abs(yr)) < 1e-12
    # the integral as the device computes it: offset + ent_q^T F, with the quadrature weights
    L.slot_q = packing.quadrature_table(L, sng)
    Q = off + packing.entry_quadrature(L) @ F
    Qr = ref.integral()
    assert np.max(np.abs(Q - Qr)) / np.max(np.abs(Qr)) < 1e-12
    # the wave segments cover every block exactly once, each inside one run of its streamed slot
    covered = np.zeros(len(L.blk_n), dtype=int)
    for g in range(L.seg_off[-1]):
        covered[L.seg_b0[g]:L.seg_b1[g]] += 1
        assert np.all(L.blk_last_slot[L.seg_b0[g]:L.seg_b1[g]] == L.seg_slot[g])
    assert np.all(covered == 1)
    assert len(L.seg_off) == packing.N_WAVES + 1


def test_vectorized_f_matches_pointwise():
    from smolyax_amd import nodes as SN
    from smolyax_amd import packing

    d, d_out = 5, 2
    k = np.log((2 + np.arange(d)) / np.log(2))
    t = OI.find_approximate_threshold(k, 150, True)
    A = np.random.default_rng(2).standard_normal((d_out, d))
    f1 = lambda x: np.cos(A @ x)
    fv = lambda X: np.cos(X @ A.T)
    blocks = packing.nonzero_blocks(k, t)
    L = packing.build_layout(blocks, SN.Leja(dim=d), d, d_out)
    F1, o1, n1 = packing.assemble_values(L, blocks, SN.Leja(dim=d), f1, {})
    F2, o2, n2 = packing.assemble_values(L, blocks, SN.Leja(dim=d), fv, {}, vectorized=True)
    assert n1 == n2
    assert np.allclose(F1, F2, rtol=0, atol=1e-15) and np.allclose(o1, o2, rtol=0, atol=1e-15)


@pytest.mark.parametrize("kind", ["leja", "gh"])
def test_f_evals_cross_compatible_with_oracle(kind):
    """The f_evals dict of the reference's format (interpolation.py:115-163, :208-228; written here
    by the oracle's set_f) is accepted as it is: every value is reused (no call of f), F and the
    offset are the ones built from f, and the dict the product writes has the oracle's keys
    (flat node keys for nested nodes, per-block node dicts for non-nested ones) and values."""
    from smolyax_amd import nodes as SN
    from smolyax_amd import packing

    d, d_out = 5, 2
    k = np.log((2 + np.arange(d)) / np.log(2))
    ong, sng = (ON.leja(dim=d), SN.Leja(dim=d)) if kind == "leja" else \
        (ON.gauss_hermite(dim=d), SN.GaussHermite(dim=d))
    t = OI.find_approximate_threshold(k, 150, ong.is_nested)
    A = np.random.default_rng(4).standard_normal((d_out, d))
    f = lambda x: np.sin(A @ x + 0.3)
    ref = OS.Interpolant(d_out=d_out, node_gen=ong, k=k, t=t)
    fe_oracle = ref.set_f(f=f)

    def no_calls(x):
        raise AssertionError("f called although f_evals holds every node")

    blocks = packing.nonzero_blocks(k, t)
    L = packing.build_layout(blocks, sng, d, d_out)
    F0, off0, _ = packing.assemble_values(L, blocks, sng, f, fe_own := {})
    F1, off1, n1 = packing.assemble_values(L, blocks, sng, no_calls, fe_oracle)
    assert n1 == 0
    assert np.array_equal(F0, F1) and np.array_equal(off0, off1)
    assert set(fe_own) == set(fe_oracle)
    for key in fe_own:
        a, b = fe_own[key], fe_oracle[key]
        if kind == "leja":
            assert np.array_equal(np.asarray(a), np.asarray(b))
        else:
            assert set(a) == set(b)
            assert all(np.array_equal(np.asarray(a[q]), np.asarray(b[q])) for q in a)

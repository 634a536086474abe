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


@pytest.mark.parametrize("kind,d,m,kscale", [("leja", 6, 200, 1.0), ("leja", 8, 3000, 0.3),
                                             ("gh", 5, 300, 1.0), ("leja", 2, 40, 1.0),
                                             ("leja", 12, 600, 0.2)])
def test_eval_program_matches_layout(kind, d, m, kscale):
    """The record program of the d_out = 1 eval kernel (records of up to three prefix factors,
    pairs of consecutive table rows, per-block path for the rest) uses every F value exactly once
    and evaluates the same interpolant as the block-by-block layout.  The index sets include
    blocks with four or more active dims (per-block path), streamed sizes outside 2..4 (whole
    segments on the per-block path) and odd prefix degrees (single records)."""
    from smolyax_amd import nodes as SN
    from smolyax_amd import packing

    rng = np.random.default_rng(d + m)
    k = kscale * np.log((2 + np.arange(d)) / np.log(2))
    sng = SN.Leja(dim=d) if kind == "leja" else SN.GaussHermite(dim=d)
    t = OI.find_approximate_threshold(k, m, sng.is_nested)
    blocks = packing.nonzero_blocks(k, t)
    L = packing.build_layout(blocks, sng, d, 1)
    A = rng.standard_normal((1, d))
    F, off, _ = packing.assemble_values(L, blocks, sng, lambda x: np.sin(A @ x + 0.3), {})
    prog = packing.build_eval_program(L)
    x = rng.uniform(-0.9, 0.9, (30, d)) * (1.0 if kind == "leja" else 2.0)
    x[0, 0] = L.slot_nodes[0, 0] if L.slot_dim[0] == 0 else x[0, 0]   # a point on a node
    y_ref = layout_eval.eval_layout(L, F, off, x)[:, 0]
    y, used = layout_eval.eval_program(L, prog, F, off, x)
    assert np.all(used == 1)
    assert np.max(np.abs(y - y_ref)) <= 1e-12 * np.max(np.abs(y_ref))
    hdr = prog.seg.reshape(-1, packing.PROG_HDR)
    # per wave: each header names the next segment's streamed dim; basis rows are the slot's
    for w in range(packing.N_WAVES):
        g0, g1 = prog.seg_off[w], prog.seg_off[w + 1]
        for g in range(g0, g1):
            ls = hdr[g, 0]
            assert hdr[g, 7] == (L.slot_dim[ls] if ls >= 0 else -1)
            if (g - g0) % packing.PROG_GROUP == 0:
                grp = hdr[g:g + packing.PROG_GROUP]
                simple = len(grp) == packing.PROG_GROUP and all(
                    h[3] == 1 and (h[4] & 0xffff) == 0 and (h[4] >> 16) == 2 for h in grp)
                assert hdr[g, 6] == int(simple)
            else:
                assert hdr[g, 6] == 0
            dl = (hdr[g, 4].view(np.uint32) >> 16) & 0xff
            bv = prog.basis.reshape(-1, packing.PROG_BASIS)[g]
            if dl:
                assert np.array_equal(bv[:dl], L.slot_nodes[ls, :dl])
                assert np.array_equal(bv[4:4 + dl], L.slot_w[ls, :dl])
    n_gen = (hdr[:, 4].view(np.uint32) >> 24).sum()
    npre = np.diff(L.blk_fac)
    print(f"{kind} d={d}: {prog.n_records} records, {n_gen} per-block blocks "
          f"(of {len(L.blk_n)}; max prefix factors {npre.max()})")

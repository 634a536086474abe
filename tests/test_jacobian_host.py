"""Dense Jacobian of the Newton form on the host, no GPU: a numpy model of what
sbi_newton_jacobian_f64 computes from the plan's tables (factor rows, adjoint entries, their CSR
over input dims) reproduces the oracle's gradient, and the singular-coordinate table gives the
oracle's NaN pattern at node coordinates.

Tolerance: 1e-11 relative to max |J| (FP64 in both; the Newton and barycentric forms differ only
in rounding)."""

import types

import numpy as np
import pytest

from oracle import indices as OI
from oracle import nodes as ON
from oracle import smolyak as OS


def jacobian_model(L, S, x):
    """J[p, o, j] = sum over the adjoint entries e of dim j of D[p, e] S[node_e, o],
    D[p, e] = h'_k(x_j) prod_{other factor rows r of e} h_r(x)."""
    P, d = x.shape
    R = L.n_tab_rows
    h, dh = np.zeros((R, P)), np.zeros((R, P))
    for i in range(L.n_tab_dims):
        j, row = L.tab_dim[i], L.tab_row[i]
        pr, dpr = np.ones(P), np.zeros(P)
        for k in range(1, L.tab_kmax[i] + 1):
            dx = x[:, j] - L.dim_xi[j, k - 1]
            dpr = dpr * dx + pr
            pr = pr * dx
            h[row + k - 1] = pr * L.dim_c[j, k]
            dh[row + k - 1] = dpr * L.dim_c[j, k]
    ent = L.vjp_ent.view(np.uint16).reshape(-1, 8)[: L.n_vjp_ent]
    node = ent[:, :2].copy().view(np.int32)[:, 0]
    rows = np.repeat(np.arange(R), np.diff(L.vjp_slot_off))
    D = dh[rows].copy()
    for f in range(6):
        r = ent[:, 2 + f].astype(np.int64)
        has = r != 0xFFFF
        D[has] *= h[r[has]]
    J = np.zeros((P, S.shape[1], d))
    for j in range(d):
        e0, e1 = L.vjp_dim_off[j], L.vjp_dim_off[j + 1]
        J[:, :, j] = D[e0:e1].T @ S[node[e0:e1]]
    # the register pass's work units: a unit covering its dim writes J, split dims add their
    # partial sums in unit order
    Ju = np.zeros_like(J)
    part = {}
    for u in L.jac_unit.reshape(-1, 4)[: L.n_jac_units]:
        i, a, b, slot = (int(v) for v in u)
        g = D[a:b].T @ S[node[a:b]]
        if slot < 0:
            Ju[:, :, L.tab_dim[i]] = g
        else:
            part[slot] = g
    for j, s0, ns in L.jac_red.reshape(-1, 3)[: L.n_jac_red]:
        Ju[:, :, j] = sum(part[s0 + q] for q in range(ns))
    assert len(part) == L.n_jac_partials
    assert np.max(np.abs(Ju - J)) <= 1e-13 * max(1.0, np.max(np.abs(J)))
    return J


def _pair(d, d_out, m, seed):
    from smolyax_amd import newton as NW
    from smolyax_amd import nodes as SN

    rng = np.random.default_rng(seed)
    k = np.log((2 + np.arange(d)) / np.log(2))
    ong, sng = ON.leja(dim=d), SN.Leja(dim=d)
    t = OI.find_approximate_threshold(k, m, True)
    A = rng.standard_normal((d_out, d)) / (np.arange(d) + 1.0)
    f = lambda x: np.sin(np.asarray(x) @ A.T + 0.3)
    ref = OS.Interpolant(d_out=d_out, node_gen=ong, k=k, t=t, f=f)
    nf = NW.NewtonForm(k, t, sng, d_out)
    zero = np.array([g(0)[0] for g in sng])
    X = np.stack([nf.grid_point(nu, zero) for nu in nf.nus])
    nf.set_values(f(X).reshape(len(X), d_out))
    return ref, nf, NW.build_newton_layout(nf), rng, k, t, sng


@pytest.mark.parametrize("d,d_out,m", [(1, 1, 9), (4, 3, 60), (8, 2, 300), (12, 1, 500),
                                       (30, 2, 3000)])
def test_newton_jacobian_tables_match_oracle(d, d_out, m):
    from smolyax_amd import newton as NW

    ref, nf, L, rng, _, _, _ = _pair(d, d_out, m, seed=d)
    x = rng.uniform(-1, 1, (40, d))
    S = np.asarray(nf.s, dtype=np.float64).reshape(len(nf.nus), d_out)
    J = jacobian_model(L, S, x)
    Jr = ref.gradient(x)
    assert rel_ok(J, Jr, 1e-11)
    # the CSR over dims covers every entry once, in order
    assert L.vjp_dim_off[0] == 0 and L.vjp_dim_off[-1] == L.n_vjp_ent
    assert np.all(np.diff(L.vjp_dim_off) >= 0)
    assert L.n_vjp_ent == sum(len(nu) for nu in nf.nus)
    # factor row of every entry and table dim of every row, as the kernels read them
    rows = np.repeat(np.arange(L.n_tab_rows), np.diff(L.vjp_slot_off))
    assert np.array_equal(L.vjp_ent_row[: L.n_vjp_ent], rows)
    for i in range(L.n_tab_dims):
        assert np.all(L.row_tdim[L.tab_row[i]: L.tab_row[i] + L.tab_kmax[i]] == i)
    units = L.jac_unit.reshape(-1, 4)[: L.n_jac_units]
    assert np.all(units[:, 2] - units[:, 1] <= NW.JAC_UNIT)
    assert np.array_equal(np.sort(np.concatenate([np.arange(a, b) for _, a, b, _ in units])),
                          np.arange(L.n_vjp_ent))
    if m >= 3000:
        assert L.n_jac_red > 0  # the heavy dims are split


def rel_ok(a, b, tol):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b))) <= tol


def test_singular_coordinates_give_the_oracle_nan_pattern():
    """x_j set to nodes of dim j of several degrees (inside and beyond the largest degree any
    block uses in dim j): NaN in J[p, :, j] exactly where x_j is in the table."""
    from smolyax_amd import packing
    from smolyax_amd.interpolation import SmolyakBarycentricInterpolator as SBI

    d, d_out = 5, 2
    ref, nf, _, rng, k, t, sng = _pair(d, d_out, 80, seed=3)
    L = packing.build_layout(packing.nonzero_blocks(k, t), sng, d, d_out)
    off, vals = SBI._singular_coordinates(types.SimpleNamespace(_layout=L, _d_in=d))
    assert off.shape == (d + 1,) and off[-1] == len(vals) or (off[-1] == 0 and len(vals) == 1)
    seq = [sng[j](12) for j in range(d)]
    x = rng.uniform(-0.9, 0.9, (36, d))
    for p in range(36):
        j, mu = p % d, p // d
        x[p, j] = seq[j][mu]
    Jr = ref.gradient(x)
    seen = set()
    for p in range(36):
        for j in range(d):
            in_table = bool(np.any((x[p, j] - vals[off[j]:off[j + 1]]) ** 2 == 0.0))
            assert np.isnan(Jr[p, :, j]).all() == in_table, (p, j)
            assert np.isnan(Jr[p, :, j]).any() == in_table, (p, j)
            if j == p % d:
                seen.add(in_table)
    assert seen == {True, False}  # node coordinates inside and beyond the table both occur


def test_singular_coordinates_gauss_hermite():
    """Non-nested nodes: every degree has its own node set.  x_j set to a node of a degree that
    some block uses in dim j gives NaN in the oracle's J[p, :, j]; a node of an unused degree does
    not; the table reproduces both."""
    from smolyax_amd import nodes as SN
    from smolyax_amd import packing
    from smolyax_amd.interpolation import SmolyakBarycentricInterpolator as SBI

    d, d_out = 3, 2
    rng = np.random.default_rng(5)
    k = np.log((2 + np.arange(d)) / np.log(2))
    ong, sng = ON.gauss_hermite(dim=d), SN.GaussHermite(dim=d)
    t = OI.find_approximate_threshold(k, 60, False)
    A = rng.standard_normal((d_out, d))
    ref = OS.Interpolant(d_out=d_out, node_gen=ong, k=k, t=t, f=lambda x: np.sin(A @ x))
    L = packing.build_layout(packing.nonzero_blocks(k, t), sng, d, d_out)
    off, vals = SBI._singular_coordinates(types.SimpleNamespace(_layout=L, _d_in=d))
    x = rng.standard_normal((40, d)) / np.sqrt(2)
    used = {}
    for dim, deg in L.slot_key:
        used.setdefault(dim, set()).add(deg)
    p = 0
    for j in range(d):
        for deg in range(0, 7):
            if p >= 40:
                break
            x[p, j] = sng[j](deg)[deg // 2]
            p += 1
    Jr = ref.gradient(x)
    seen = set()
    for q in range(40):
        for j in range(d):
            in_table = bool(np.any((x[q, j] - vals[off[j]:off[j + 1]]) ** 2 == 0.0))
            assert np.isnan(Jr[q, :, j]).all() == in_table, (q, j)
            assert np.isnan(Jr[q, :, j]).any() == in_table, (q, j)
            seen.add(in_table)
    assert seen == {True, False}

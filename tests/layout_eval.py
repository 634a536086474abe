"""Numpy evaluation of the packed device layout (test helper, CPU).

Mirrors what the HIP kernels compute on `smolyax_amd.packing.Layout` so that the host packing
(slot tables, block order, F rows, zeta scaling) can be checked against the oracle without a GPU.
"""

import itertools as it

import numpy as np


def first_form_basis(x, xi, w):
    """(P,) x, (m,) nodes/weights -> (P, m) Lagrange basis, exact one-hot at nodes."""
    d = x[:, None] - xi[None, :]
    m = len(xi)
    out = np.empty((len(x), m))
    for a in range(m):
        p = np.full(len(x), w[a])
        for q in range(m):
            if q != a:
                p = p * d[:, q]
        out[:, a] = p
    hit = d == 0
    rows = hit.any(axis=1)
    out[rows] = hit[rows].astype(float)
    return out


def eval_layout(L, F, offset, x):
    P = x.shape[0]
    y = np.tile(offset, (P, 1)).astype(float)
    for b in range(len(L.blk_n)):
        fd, fg = L.blk_dims[b], L.blk_degs[b]
        slots = [L.slot_key.index((int(a), int(g))) for a, g in zip(fd, fg)]
        bases = [first_form_basis(x[:, a], L.slot_nodes[s, : g + 1], L.slot_w[s, : g + 1])
                 for a, g, s in zip(fd, fg, slots)]
        r0 = L.blk_foff[b]
        for e, mu in enumerate(it.product(*[range(g + 1) for g in fg])):
            wgt = np.ones(P)
            for i, m in enumerate(mu):
                wgt = wgt * bases[i][:, m]
            y += wgt[:, None] * F[r0 + e][None, :]
    return y

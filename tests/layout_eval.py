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


def eval_program(L, prog, F, offset, x):
    """The record program of the d_out = 1 eval kernel on the CPU: per program segment the
    records' V[m] += (product of table values) * F, then y += V . (streamed basis); per-block
    blocks entry by entry.  Also returns how often every F position was used."""
    from smolyax_amd import packing

    P = x.shape[0]
    Ff = F.reshape(-1)
    pF = packing.program_values(prog, F)
    used = np.zeros(len(Ff), dtype=np.int64)
    tab = np.zeros((prog.ones_row + 2, P))
    tab[prog.ones_row:] = 1.0
    for s, r0 in zip(L.prefix_slot, L.prefix_off):
        g = int(L.slot_deg[s])
        tab[r0:r0 + g + 1] = first_form_basis(x[:, L.slot_dim[s]], L.slot_nodes[s, : g + 1],
                                              L.slot_w[s, : g + 1]).T
    y = np.full(P, float(offset.reshape(-1)[0]))
    hdr = prog.seg.view(np.uint32).reshape(-1, packing.PROG_HDR).astype(np.int64)
    hdr[:, 0] = prog.seg.reshape(-1, packing.PROG_HDR)[:, 0]
    for h in hdr:
        ls, f_off, r_off = int(h[0]), int(h[1]), int(h[2])
        counts = [h[3] & 0xffff, h[3] >> 16, h[4] & 0xffff]
        dl, n_gen, gen_b0 = int((h[4] >> 16) & 0xff), int(h[4] >> 24), int(h[5])
        if ls >= 0:
            g = int(L.slot_deg[ls])
            Ll = first_form_basis(x[:, L.slot_dim[ls]], L.slot_nodes[ls, : g + 1],
                                  L.slot_w[ls, : g + 1])
        else:
            Ll = np.ones((P, 1))
        if dl:
            V = np.zeros((P, dl))
            for npre, n in zip(packing.PROG_TYPES, counts):
                for _ in range(int(n)):
                    rr = prog.rows[r_off:r_off + npre]
                    fi = prog.f_index[f_off:f_off + 2 * dl]
                    fv = pF[f_off:f_off + 2 * dl]
                    r_off += npre
                    f_off += 2 * dl
                    used[fi[fi >= 0]] += 1
                    pre = np.ones(P)
                    for r in rr[:-1]:
                        pre = pre * tab[r]
                    for v in range(2):
                        V += (pre * tab[rr[-1] + v])[:, None] * fv[None, v * dl:(v + 1) * dl]
            y += np.sum(V * Ll[:, :dl], axis=1)
        for b in range(gen_b0, gen_b0 + n_gen):
            fd, fg = L.blk_dims[b], L.blk_degs[b]
            slots = [L.slot_key.index((int(a), int(gg))) for a, gg in zip(fd, fg)]
            bases = [first_form_basis(x[:, a], L.slot_nodes[s, : gg + 1], L.slot_w[s, : gg + 1])
                     for a, gg, s in zip(fd, fg, slots)]
            r0 = int(L.blk_foff[b])
            for e, mu in enumerate(it.product(*[range(gg + 1) for gg in fg])):
                wgt = np.ones(P)
                for i, m in enumerate(mu):
                    wgt = wgt * bases[i][:, m]
                y += wgt * Ff[r0 + e]
                used[r0 + e] += 1
    return y, used

"""Host-side packing of a Smolyak interpolant into the device layout (numpy only, no GPU).

Replaces set_f's grouped, zero-padded layout (src/smolyax/interpolation.py:168-235) with a ragged
layout: every block (multi-index nu != 0 with zeta_nu != 0) keeps its own shape prod(nu_j + 1),
so no padded entries are stored or multiplied.

Per block the factors are the active dims in ascending order.  The last factor is "streamed"
(its 1D basis is computed per run in registers) when its degree is <= 11, otherwise the block has
no streamed factor and all its factors are "prefix" factors.  Prefix slots get rows of an on-chip
basis table.  Blocks are sorted by streamed slot (dim, deg) so runs share one streamed basis.
"""

import itertools as it
from dataclasses import dataclass, field

import numpy as np

from . import barycentric, indices

STREAM_LMAX = 12     # max (deg + 1) of a streamed factor (kLMax in bary.hip)
SLOT_ROW_MIN = 12    # slot rows are padded to at least this many entries
N_WAVES = 16         # waves per point tile (kWaves in bary.hip)
# Wave split of the barycentric eval kernel, in F entries: a block costs its entries plus
# BLOCK_COST, a run start SEG_COST more (the streamed basis and the run's metadata are fetched
# anew); tuned on the d_in = 1000 workload, where balancing by entries alone left the last wave
# with 490 runs and twice the run time.
BLOCK_COST = 4.0
SEG_COST = 35.0


@dataclass
class Blocks:
    """Blocks in discovery order (as produced by the index machinery)."""
    offsets: np.ndarray   # CSR into dims / degs
    dims: np.ndarray
    degs: np.ndarray
    zetas: np.ndarray
    zeta0: int            # coefficient of nu = 0 (0 if it vanishes)


def nonzero_blocks(k, t) -> Blocks:
    off, dims, degs, zet = indices.nonzero_blocks(k, t)
    zeta0 = 0
    keep = []
    for i in range(len(zet)):
        if off[i + 1] == off[i]:
            zeta0 = int(zet[i])
        else:
            keep.append(i)
    keep = np.asarray(keep, dtype=np.int64)
    lens = off[keep + 1] - off[keep]
    noff = np.zeros(len(keep) + 1, dtype=np.int64)
    noff[1:] = np.cumsum(lens)
    idx = np.concatenate([np.arange(off[i], off[i + 1]) for i in keep]) if len(keep) else np.zeros(0, np.int64)
    return Blocks(noff, dims[idx].astype(np.int32), degs[idx].astype(np.int32),
                  zet[keep].astype(np.int64), zeta0)


@dataclass
class Layout:
    d_in: int
    d_out: int
    nested: bool
    # slots
    slot_key: list = field(default_factory=list)      # (dim, deg)
    slot_dim: np.ndarray = None
    slot_deg: np.ndarray = None
    slot_nodes: np.ndarray = None                     # (S, dmax1)
    slot_w: np.ndarray = None                         # first-form weights
    slot_q: np.ndarray = None                         # quadrature weights (filled lazily)
    dmax1: int = SLOT_ROW_MIN
    # prefix table
    prefix_slot: np.ndarray = None
    prefix_off: np.ndarray = None
    n_prefix_vals: int = 0
    # blocks in device order
    order: np.ndarray = None                          # device block b -> discovery index
    blk_n: np.ndarray = None
    blk_fac: np.ndarray = None
    fac_row: np.ndarray = None
    fac_deg: np.ndarray = None
    fac_slot: np.ndarray = None
    blk_last_slot: np.ndarray = None
    blk_foff: np.ndarray = None
    blk_zeta: np.ndarray = None
    blk_dims: list = None                             # factor dims per device block
    blk_degs: list = None
    run_slot: np.ndarray = None
    run_start: np.ndarray = None
    blk_desc: np.ndarray = None
    seg_off: np.ndarray = None
    seg_slot: np.ndarray = None
    seg_b0: np.ndarray = None
    seg_b1: np.ndarray = None
    dim_blk: np.ndarray = None
    dim_blk_idx: np.ndarray = None
    dim_blk_pos: np.ndarray = None
    k_raw: int = 0
    max_n: int = 0
    zeta0: int = 0


def build_layout(blocks: Blocks, node_gen, d_in: int, d_out: int) -> Layout:
    L = Layout(d_in=d_in, d_out=d_out, nested=node_gen.is_nested, zeta0=blocks.zeta0)
    nb = len(blocks.zetas)
    off, dims, degs = blocks.offsets, blocks.dims, blocks.degs

    slot_id = {}

    def slot(dim, deg):
        key = (int(dim), int(deg))
        s = slot_id.get(key)
        if s is None:
            s = len(L.slot_key)
            slot_id[key] = s
            L.slot_key.append(key)
        return s

    last_slot = np.full(nb, -1, dtype=np.int64)
    pre_slots = []
    for b in range(nb):
        fd, fg = dims[off[b]:off[b + 1]], degs[off[b]:off[b + 1]]
        sl = [slot(a, g) for a, g in zip(fd, fg)]
        if fg[-1] + 1 <= STREAM_LMAX:
            last_slot[b] = sl[-1]
            pre_slots.append(sl[:-1])
        else:
            pre_slots.append(sl)

    S = len(L.slot_key)
    L.slot_dim = np.array([k[0] for k in L.slot_key], dtype=np.int32)
    L.slot_deg = np.array([k[1] for k in L.slot_key], dtype=np.int32)
    maxdeg = int(L.slot_deg.max()) if S else 0
    L.dmax1 = max(SLOT_ROW_MIN, maxdeg + 1)
    L.slot_nodes = np.zeros((max(S, 1), L.dmax1))
    L.slot_w = np.zeros((max(S, 1), L.dmax1))
    cache = {}
    for s, (dim, deg) in enumerate(L.slot_key):
        g = node_gen[dim]
        ck = (g._table_key(), deg)
        v = cache.get(ck)
        if v is None:
            xi = np.asarray(g(deg), dtype=np.float64)
            v = (xi, barycentric.first_form_weights(xi))
            cache[ck] = v
        L.slot_nodes[s, : deg + 1] = v[0]
        L.slot_w[s, : deg + 1] = v[1]

    # prefix slots / table rows
    used_pre = sorted({s for ps in pre_slots for s in ps})
    row_of = {}
    rows = 0
    for s in used_pre:
        row_of[s] = rows
        rows += int(L.slot_deg[s]) + 1
    L.prefix_slot = np.array(used_pre, dtype=np.int32)
    L.prefix_off = np.array([row_of[s] for s in used_pre], dtype=np.int32)
    L.n_prefix_vals = rows

    # device order: runs of equal streamed slot, streamed dims ascending (no-stream blocks first)
    def run_key(b):
        s = last_slot[b]
        return (-1, -1) if s < 0 else (int(L.slot_dim[s]), int(L.slot_deg[s]))

    order = sorted(range(nb), key=run_key)
    L.order = np.array(order, dtype=np.int64)
    blk_n, blk_fac, fac_row, fac_deg, fac_slot, blk_last, foff, zeta = [], [0], [], [], [], [], [0], []
    L.blk_dims, L.blk_degs = [], []
    run_slot, run_start = [], []
    for pos, b in enumerate(order):
        fd, fg = dims[off[b]:off[b + 1]], degs[off[b]:off[b + 1]]
        blk_n.append(len(fd))
        for s in pre_slots[b]:
            fac_row.append(row_of[s])
            fac_deg.append(int(L.slot_deg[s]))
            fac_slot.append(s)
        blk_fac.append(len(fac_row))
        blk_last.append(int(last_slot[b]))
        foff.append(foff[-1] + int(np.prod(fg.astype(np.int64) + 1)))
        zeta.append(int(blocks.zetas[b]))
        L.blk_dims.append(fd)
        L.blk_degs.append(fg)
        if not run_slot or run_slot[-1] != last_slot[b]:
            run_slot.append(int(last_slot[b]))
            run_start.append(pos)
    run_start.append(nb)
    L.blk_n = np.array(blk_n, dtype=np.int32)
    L.blk_fac = np.array(blk_fac, dtype=np.int32)
    L.fac_row = np.array(fac_row or [0], dtype=np.int32)
    L.fac_deg = np.array(fac_deg or [0], dtype=np.int32)
    L.fac_slot = np.array(fac_slot or [0], dtype=np.int32)
    L.blk_last_slot = np.array(blk_last or [0], dtype=np.int32)
    L.blk_foff = np.array(foff, dtype=np.int64)
    L.blk_zeta = np.array(zeta, dtype=np.int64)
    L.run_slot = np.array(run_slot, dtype=np.int32)
    L.run_start = np.array(run_start, dtype=np.int32)
    L.k_raw = int(foff[-1])
    L.max_n = int(max(blk_n)) if blk_n else 0

    # packed per-block records for the eval kernel
    desc = np.zeros((max(nb, 1), 4), dtype=np.int64)
    for pos in range(nb):
        f0, f1 = L.blk_fac[pos], L.blk_fac[pos + 1]
        npre = f1 - f0
        ls = L.blk_last_slot[pos]
        dl = int(L.slot_deg[ls]) if ls >= 0 else 0
        dg = [int(L.fac_deg[f0 + i]) for i in range(min(npre, 3))] + [0] * (3 - min(npre, 3))
        r = [int(L.fac_row[f0 + i]) for i in range(min(npre, 3))] + [0] * (3 - min(npre, 3))
        assert dl < STREAM_LMAX
        if npre > 3 or max(dg) >= 32 or max(r) >= 65536:
            packed = 7 | (dl << 3)  # generic record: the kernel reads fac_deg / fac_row
            r = [0, 0, 0]
        else:
            packed = npre | (dl << 3) | (dg[0] << 8) | (dg[1] << 13) | (dg[2] << 18)
        assert L.blk_foff[pos] < 2**31 and ls + 1 < 65536
        desc[pos] = [L.blk_foff[pos], packed, r[0] | (r[1] << 16), r[2] | ((ls + 1) << 16)]
    L.blk_desc = desc.astype(np.uint32).view(np.int32).reshape(-1)

    # cost-balanced split of the blocks over the waves of a tile; segments = runs cut at the
    # wave boundaries (a wave recomputes the streamed basis at each segment start, so a run start
    # costs SEG_COST entries more)
    bcost = np.diff(L.blk_foff).astype(np.float64) + BLOCK_COST
    if nb:
        bcost[L.run_start[:-1][L.run_start[:-1] < nb]] += SEG_COST
    cum = np.concatenate([[0.0], np.cumsum(bcost)]) if nb else np.zeros(1)
    cuts = [0]
    for w in range(1, N_WAVES):
        cuts.append(int(np.searchsorted(cum, cum[-1] * w / N_WAVES, side="left")))
    cuts.append(nb)
    cuts = np.maximum.accumulate(np.clip(np.array(cuts), 0, nb))
    seg_slot, seg_b0, seg_b1, seg_off = [], [], [], [0]
    for w in range(N_WAVES):
        b0, b1 = int(cuts[w]), int(cuts[w + 1])
        b = b0
        while b < b1:
            r = int(np.searchsorted(L.run_start, b, side="right")) - 1
            e = min(int(L.run_start[r + 1]), b1)
            seg_slot.append(int(L.run_slot[r]))
            seg_b0.append(b)
            seg_b1.append(e)
            b = e
        seg_off.append(len(seg_slot))
    L.seg_off = np.array(seg_off, dtype=np.int32)
    L.seg_slot = np.array(seg_slot or [0], dtype=np.int32)
    L.seg_b0 = np.array(seg_b0 or [0], dtype=np.int32)
    L.seg_b1 = np.array(seg_b1 or [0], dtype=np.int32)

    # dim -> (block, factor position) CSR for the jacobian
    lists = [[] for _ in range(d_in)]
    for pos, b in enumerate(order):
        # factors are in ascending dim order on the device too (prefix factors, then the
        # streamed one, which is the highest dim), so the position is the rank of the dim
        for i, dm in enumerate(dims[off[b]:off[b + 1]]):
            lists[int(dm)].append((pos, i))
    L.dim_blk = np.zeros(d_in + 1, dtype=np.int32)
    L.dim_blk[1:] = np.cumsum([len(v) for v in lists])
    flat = [e for v in lists for e in v]
    L.dim_blk_idx = np.array([e[0] for e in flat] or [0], dtype=np.int32)
    L.dim_blk_pos = np.array([e[1] for e in flat] or [0], dtype=np.int32)
    return L


def quadrature_table(L: Layout, node_gen) -> np.ndarray:
    q = np.zeros_like(L.slot_nodes)
    for s, (dim, deg) in enumerate(L.slot_key):
        w = np.asarray(node_gen[dim].get_quadrature_weights(deg), dtype=np.float64)
        q[s, : len(w)] = w
    return q


def entry_quadrature(L: Layout) -> np.ndarray:
    """Quadrature weight of every entry of F in device order: the outer product of the block's
    factor weights (prefix factors, then the streamed one), C order, as in F
    (interpolation.py:347-390 forms the same products per group)."""
    out = np.empty(L.k_raw, dtype=np.float64)
    for b in range(len(L.blk_n)):
        slots = list(L.fac_slot[L.blk_fac[b]:L.blk_fac[b + 1]])
        if L.blk_last_slot[b] >= 0:
            slots.append(L.blk_last_slot[b])
        w = np.ones(1)
        for s in slots:
            w = np.multiply.outer(w, L.slot_q[s, : L.slot_deg[s] + 1]).reshape(-1)
        out[L.blk_foff[b]:L.blk_foff[b + 1]] = w
    return out


def entry_keys(fd, fg):
    """Grid entries of one block in C order (last factor fastest): (mu tuple, sparse key)."""
    n = len(fd)
    for mu in it.product(*[range(int(g) + 1) for g in fg]):
        key = tuple((int(fd[i]), mu[i]) for i in range(n) if mu[i] > 0)
        yield mu, key


def assemble_values(L: Layout, blocks: Blocks, node_gen, f, f_evals, vectorized=False):
    """Evaluate f at the grid nodes (reusing f_evals) and fill the zeta-scaled device F.

    f_evals follows the reference format (interpolation.py:141-228): for nested nodes a flat dict
    keyed by sparse node tuples ((dim, mu), ...) (mu > 0, dims ascending); for non-nested nodes a
    dict keyed by the block's sparse multi-index, each holding such a node dict.
    Returns (F[k_raw, d_out], offset[d_out], n_new).
    """
    d_in, d_out = L.d_in, L.d_out
    zero = np.array([g(0)[0] for g in node_gen], dtype=np.float64)
    n_new = 0
    pending = []  # (dict, key, x) for vectorized evaluation

    def get(store, key, coords):
        nonlocal n_new
        if key in store:
            return store[key]
        x = zero.copy()
        for dm, val in coords:
            x[dm] = val
        if vectorized:
            pending.append((store, key, x))
            store[key] = None
            n_new += 1
            return None
        v = f(x)
        store[key] = v
        n_new += 1
        return v

    # offset term (nu = 0), evaluated only when its coefficient is non-zero (interpolation.py:151)
    root_store = f_evals if L.nested else f_evals.setdefault((), {})
    if L.zeta0 != 0:
        get(root_store, (), [])

    # grid node values per slot
    node_vals = [L.slot_nodes[s, : L.slot_deg[s] + 1] for s in range(len(L.slot_key))]
    slot_of = {k: s for s, k in enumerate(L.slot_key)}
    entries = []  # (store, key) per F row in device order
    for pos, b in enumerate(L.order):
        fd, fg = L.blk_dims[pos], L.blk_degs[pos]
        nu = tuple((int(a), int(g)) for a, g in zip(fd, fg))
        store = f_evals if L.nested else f_evals.setdefault(nu, {})
        sl = [slot_of[(int(a), int(g))] for a, g in zip(fd, fg)]
        for mu, key in entry_keys(fd, fg):
            coords = [(int(fd[i]), node_vals[sl[i]][mu[i]]) for i in range(len(fd))]
            get(store, key, coords)
            entries.append((store, key))

    if vectorized and pending:
        X = np.stack([p[2] for p in pending])
        Y = np.asarray(f(X), dtype=np.float64).reshape(len(pending), -1)
        for (store, key, _), yv in zip(pending, Y):
            store[key] = yv

    F = np.empty((L.k_raw, d_out), dtype=np.float64)
    zrow = np.repeat(L.blk_zeta, np.diff(L.blk_foff)).astype(np.float64)
    for r, (store, key) in enumerate(entries):
        F[r] = np.broadcast_to(np.asarray(store[key], dtype=np.float64).reshape(-1), (d_out,))
    F *= zrow[:, None]
    if L.zeta0 != 0:
        v = np.broadcast_to(np.asarray(root_store[()], dtype=np.float64).reshape(-1), (d_out,))
        offset = float(L.zeta0) * v
    else:
        offset = np.zeros(d_out)
    return F, np.ascontiguousarray(offset, dtype=np.float64), n_new

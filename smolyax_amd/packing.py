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
# Record program of the d_out = 1 eval kernel: streamed sizes (deg + 1) with record code
# (kProgDL in bary.hip), the record types in program order (prefix factors per record), the
# header size of a program segment in int32, and the table rows the kernel holds in LDS.
PROG_DL = (2, 3, 4)
PROG_MAX_PRE = 3
PROG_TYPES = (1, 2, 3)
PROG_HDR = 8
PROG_BASIS = 8           # doubles per segment: nodes and weights of its streamed slot (DL <= 4)
PROG_GROUP = 4           # segments per group of the eval kernel (kXGroup in bary.hip)
PROG_MAX_GEN = 255       # per-block (generic) blocks per program segment
PROG_MAX_COUNT = 65535   # records of one type per program segment
PROG_MAX_ROWS = 320      # prefix table rows + 2 rows of ones, 64 points x 8 B each, in 160 KiB
PROG_SGPR = 48           # SGPRs per record chunk (kProgSgpr in bary.hip)
# Wave split of the program: per-wave cost in cycles as measured on MI355X at d_in = 1000 (one
# scalar memory round trip per record chunk): a full chunk, a partial chunk, a segment start
# beyond its chunks, a block on the per-block path per entry (plus a segment-like start).
PROG_CYC_FULL = 790.0
PROG_CYC_TAIL = 690.0
PROG_CYC_SEG = 1390.0
PROG_CYC_GEN_ENTRY = 100.0
PROG_CYC_GROUP1 = 2400.0   # a group of PROG_GROUP one-record segments (one chunk for all)


def prog_chunk(dl: int, npre: int) -> int:
    """Records per chunk of the eval kernel for a streamed size and record type."""
    return max(1, min(12, PROG_SGPR // (4 * dl + npre)))


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
        # within a run, blocks with more than PROG_MAX_PRE prefix factors come last: the record
        # program of the eval kernel leaves them to its per-block path, which takes a segment's
        # trailing blocks
        s = last_slot[b]
        many = len(pre_slots[b]) > PROG_MAX_PRE
        return (-1, -1, many) if s < 0 else (int(L.slot_dim[s]), int(L.slot_deg[s]), many)

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


@dataclass
class ValueTable:
    """f at the distinct grid nodes and the map from F's rows to them:
    F[r] = row_scale[r] * V[row_node[r]] (the block's zeta folded in)."""
    V: np.ndarray            # (n_nodes, d_out) f values, one row per distinct grid node
    row_node: np.ndarray     # (k_raw,) int32 row of V for each row of F
    row_scale: np.ndarray    # (k_raw,) zeta of the row's block
    offset: np.ndarray       # (d_out,) zeta_0 * f(root)
    n_new: int               # calls of f (new nodes)


def stack_values(vals, d_out):
    """Rows of f values (scalars or arrays of d_out, or of 1 to be broadcast) as (n, d_out)."""
    n = len(vals)
    if n == 0:
        return np.zeros((0, d_out))
    try:
        A = np.asarray(vals, dtype=np.float64).reshape(n, -1)
    except ValueError:  # ragged: mixed scalars and arrays
        A = np.stack([np.asarray(v, dtype=np.float64).reshape(-1) for v in vals]) \
            if all(np.size(v) == d_out for v in vals) else None
        if A is None:
            A = np.empty((n, d_out))
            for i, v in enumerate(vals):
                A[i] = np.broadcast_to(np.asarray(v, dtype=np.float64).reshape(-1), (d_out,))
    if A.shape[1] != d_out:
        A = np.ascontiguousarray(np.broadcast_to(A, (n, d_out)))
    return A


def gather_values(L: Layout, blocks: Blocks, node_gen, f, f_evals, vectorized=False) -> ValueTable:
    """Evaluate f at the grid nodes (reusing f_evals) and map F's rows to the distinct nodes.

    f_evals follows the reference format (interpolation.py:141-228): for nested nodes a flat dict
    keyed by sparse node tuples ((dim, mu), ...) (mu > 0, dims ascending); for non-nested nodes a
    dict keyed by the block's sparse multi-index, each holding such a node dict.  For nested
    nodes V has one row per distinct grid node (blocks share them); for non-nested ones one row
    per row of F.
    """
    d_in, d_out = L.d_in, L.d_out
    zero = np.array([g(0)[0] for g in node_gen], dtype=np.float64)
    n_new = 0
    pending = []  # (dict, key, x) for vectorized evaluation

    def get(store, key, coords):
        nonlocal n_new
        if key in store:
            return
        x = zero.copy()
        for dm, val in coords:
            x[dm] = val
        if vectorized:
            pending.append((store, key, x))
            store[key] = None
        else:
            store[key] = f(x)
        n_new += 1

    # offset term (nu = 0), evaluated only when its coefficient is non-zero (interpolation.py:151)
    root_store = f_evals if L.nested else f_evals.setdefault((), {})
    if L.zeta0 != 0:
        get(root_store, (), [])

    node_vals = [L.slot_nodes[s, : L.slot_deg[s] + 1] for s in range(len(L.slot_key))]
    slot_of = {k: s for s, k in enumerate(L.slot_key)}
    row_node = np.empty(L.k_raw, dtype=np.int32)
    uid = {}           # nested: node key -> row of V
    src = []           # (store, key) per row of V
    r = 0
    for pos, b in enumerate(L.order):
        fd, fg = L.blk_dims[pos], L.blk_degs[pos]
        nu = tuple((int(a), int(g)) for a, g in zip(fd, fg))
        store = f_evals if L.nested else f_evals.setdefault(nu, {})
        sl = [slot_of[(int(a), int(g))] for a, g in zip(fd, fg)]
        for mu, key in entry_keys(fd, fg):
            if key not in store:
                get(store, key, [(int(fd[i]), node_vals[sl[i]][mu[i]]) for i in range(len(fd))])
            if L.nested:
                u = uid.get(key)
                if u is None:
                    u = uid[key] = len(src)
                    src.append((store, key))
            else:
                u = len(src)
                src.append((store, key))
            row_node[r] = u
            r += 1

    if vectorized and pending:
        X = np.stack([p[2] for p in pending])
        Y = np.asarray(f(X), dtype=np.float64).reshape(len(pending), -1)
        for (store, key, _), yv in zip(pending, Y):
            store[key] = yv

    V = stack_values([st[key] for st, key in src], d_out)
    row_scale = np.repeat(L.blk_zeta, np.diff(L.blk_foff)).astype(np.float64)
    if L.zeta0 != 0:
        v = np.broadcast_to(np.asarray(root_store[()], dtype=np.float64).reshape(-1), (d_out,))
        offset = float(L.zeta0) * v
    else:
        offset = np.zeros(d_out)
    return ValueTable(V=V, row_node=row_node, row_scale=row_scale,
                      offset=np.ascontiguousarray(offset, dtype=np.float64), n_new=n_new)


def pack_rows(V, row_node, row_scale):
    """Host form of the device packing (sbi_pack_rows_f64): F[r] = row_scale[r] * V[row_node[r]],
    zero where row_node is -1."""
    Vz = np.concatenate([np.asarray(V, dtype=np.float64), np.zeros((1, V.shape[1]))])
    return Vz[row_node] * np.asarray(row_scale, dtype=np.float64)[:, None]


def assemble_values(L: Layout, blocks: Blocks, node_gen, f, f_evals, vectorized=False):
    """The zeta-scaled F on the host (the device builds the same from gather_values' table with
    sbi_pack_rows_f64).  Returns (F[k_raw, d_out], offset[d_out], n_new)."""
    vt = gather_values(L, blocks, node_gen, f, f_evals, vectorized)
    return pack_rows(vt.V, vt.row_node, vt.row_scale), vt.offset, vt.n_new


@dataclass
class EvalProgram:
    """Record program of the d_out = 1 barycentric eval kernel (see build_eval_program)."""
    seg_off: np.ndarray     # [N_WAVES + 1] per wave: range of its program segments
    seg: np.ndarray         # [n_seg * PROG_HDR] int32 segment headers
    basis: np.ndarray       # [n_seg * PROG_BASIS] nodes and weights of each segment's streamed slot
    rows: np.ndarray        # int32 table rows, record order
    f_index: np.ndarray     # int64: position in the flat device F of every program F value (-1: 0)
    n_records: int
    ones_row: int           # first of the two table rows that hold 1.0


def _block_records(L: Layout, pos: int, dl: int, ones: int):
    """Records of one block with at most PROG_MAX_PRE prefix factors: yields (prefix factors,
    rows, F positions).  A record fixes every prefix factor but the last and two consecutive values
    c0, c0 + 1 of the last one (rows r + c0 and r + c0 + 1 of the table); the block's F is C order
    over (prefix factors..., streamed factor), so the record's 2 DL values are contiguous.  An odd
    last value gets zeros as its second half (the row after it is still a table row: the next
    slot's, or a row of ones).  A block without prefix factors is one record on the rows of ones
    with zeros as its second half."""
    f0, f1 = int(L.blk_fac[pos]), int(L.blk_fac[pos + 1])
    npre = f1 - f0
    base = int(L.blk_foff[pos])
    zeros = [-1] * dl
    if npre == 0:
        yield 1, [ones], list(range(base, base + dl)) + zeros
        return
    rows = [int(r) for r in L.fac_row[f0:f1]]
    degs = [int(g) for g in L.fac_deg[f0:f1]]
    strides = [0] * npre
    st = dl
    for i in range(npre - 1, -1, -1):
        strides[i] = st
        st *= degs[i] + 1
    for outer in it.product(*[range(g + 1) for g in degs[:-1]]):
        o_off = base + sum(m * strides[i] for i, m in enumerate(outer))
        o_rows = [rows[i] + m for i, m in enumerate(outer)]
        for c0 in range(0, degs[-1] + 1, 2):
            a = o_off + c0 * dl
            second = list(range(a + dl, a + 2 * dl)) if c0 + 1 <= degs[-1] else zeros
            yield npre, o_rows + [rows[-1] + c0], list(range(a, a + dl)) + second


def build_eval_program(L: Layout):
    """Record program of the d_out = 1 eval kernel; None when the table with its rows of ones
    exceeds PROG_MAX_ROWS.

    The blocks (device order) are split over the N_WAVES waves of a tile by cost (PROG_* costs);
    a program segment is the part of one run (one streamed slot) inside one wave, cut further to
    the header's count limits.  Its records come grouped by their number of prefix factors
    (PROG_TYPES order), then (per-block path) its trailing blocks with more than PROG_MAX_PRE
    prefix factors, or all of its blocks when the streamed size is not in PROG_DL or there is no
    streamed factor.  Header (int32):
      [0] streamed slot (-1: none)   [1] first F value (program F)   [2] first row
      [3] n(1) | n(2) << 16          [4] n(3) | DL << 16 | n_gen << 24 (DL = 0: no records)
      [5] first per-block block      [6] 1 on the first segment of a group of PROG_GROUP
      segments (counted from the wave's first) that have one record each, DL = 2 and nothing
      on the per-block path, else 0   [7] streamed dim of this segment (-1: none)
    Basis (double, PROG_BASIS per segment): nodes then first-form weights of the streamed slot.
    """
    ones = int(L.n_prefix_vals)
    if ones + 2 > PROG_MAX_ROWS:
        return None
    nt = len(PROG_TYPES)
    nb = len(L.blk_n)
    recs_of = []
    cnt = np.zeros((nb, nt), dtype=np.int64)      # records of each type per block
    gen_e = np.zeros(nb)                          # entries of a per-block-path block
    run_of = np.zeros(nb, dtype=np.int64)
    for r in range(len(L.run_start) - 1):
        run_of[int(L.run_start[r]):int(L.run_start[r + 1])] = r
    for pos in range(nb):
        ls = int(L.blk_last_slot[pos])
        dl = int(L.slot_deg[ls]) + 1 if ls >= 0 else 0
        npre = int(L.blk_fac[pos + 1] - L.blk_fac[pos])
        if dl in PROG_DL and npre <= PROG_MAX_PRE:
            r = list(_block_records(L, pos, dl, ones))
            for tp, _, _ in r:
                cnt[pos, PROG_TYPES.index(tp)] += 1
        else:
            r = None
            gen_e[pos] = float(L.blk_foff[pos + 1] - L.blk_foff[pos])
        recs_of.append(r)
    units = []
    for pos in range(nb):
        ls = int(L.blk_last_slot[pos])
        dl = int(L.slot_deg[ls]) + 1 if ls >= 0 else 0
        units.append([prog_chunk(dl, tp) if dl in PROG_DL else 1 for tp in PROG_TYPES])

    def seg_cost(c, u, ge):
        cyc = PROG_CYC_SEG + ge * PROG_CYC_GEN_ENTRY
        for i in range(nt):
            if c[i]:
                cyc += (c[i] // u[i]) * PROG_CYC_FULL + (PROG_CYC_TAIL if c[i] % u[i] else 0.0)
        return cyc

    dl_of = [int(L.slot_deg[int(L.blk_last_slot[pos])]) + 1 if L.blk_last_slot[pos] >= 0 else 0
             for pos in range(nb)]

    def group_cost(grp):
        if len(grp) == PROG_GROUP and all(one for _, one in grp):
            return PROG_CYC_GROUP1
        return sum(cyc for cyc, _ in grp)

    cntl, gel, runl = cnt.tolist(), gen_e.tolist(), run_of.tolist()

    def split(bound):
        """Greedy cut of the block sequence into waves of cost <= bound (segments in groups of
        PROG_GROUP from each wave's first, as the kernel walks them); the cut list or None.
        split.last is the cost of the last wave."""
        cuts = [0]
        done, grp = 0.0, []                # cost of the wave's complete groups; the open group
        seg_c, seg_ge, run = [0] * nt, 0.0, -1
        for pos in range(nb):
            new_seg = runl[pos] != run
            c = list(cntl[pos]) if new_seg else [a + b for a, b in zip(seg_c, cntl[pos])]
            ge = (0.0 if new_seg else seg_ge) + gel[pos]
            info = (seg_cost(c, units[pos], ge), dl_of[pos] == 2 and c == [1, 0, 0] and ge == 0)
            if new_seg:
                d2, g2 = (done + group_cost(grp), [info]) if len(grp) == PROG_GROUP else \
                    (done, grp + [info])
            else:
                d2, g2 = done, grp[:-1] + [info]
            if d2 + group_cost(g2) > bound and pos > cuts[-1]:
                cuts.append(pos)
                if len(cuts) > N_WAVES:
                    return None
                c, ge = list(cntl[pos]), gel[pos]
                info = (seg_cost(c, units[pos], ge), dl_of[pos] == 2 and c == [1, 0, 0] and ge == 0)
                d2, g2 = 0.0, [info]
            done, grp = d2, g2
            seg_c, seg_ge, run = c, ge, runl[pos]
        split.last = done + group_cost(grp)
        return cuts + [nb] * (N_WAVES + 1 - len(cuts))

    # min-max bound by bisection, bracketed around an even share of the one-wave cost and
    # narrowed to 2e-4 of it (a fraction of one record chunk's cost per wave)
    split(float("inf"))
    share = max(split.last / N_WAVES, 1.0)
    lo, hi = 0.9 * share, 1.1 * share
    while split(hi) is None:
        hi *= 1.25
    while lo > 1e-9 and split(lo) is not None:
        lo *= 0.5
    while hi - lo > 2e-4 * hi:
        mid = 0.5 * (lo + hi)
        if split(mid) is None:
            lo = mid
        else:
            hi = mid
    cuts = np.array(split(hi))
    seg_off, hdrs, basis, rows, fidx = [0], [], [], [], []
    nrec = 0
    for w in range(N_WAVES):
        w_hdrs = []
        b, b_end = int(cuts[w]), int(cuts[w + 1])
        while b < b_end:
            r = int(np.searchsorted(L.run_start, b, side="right")) - 1
            e = min(int(L.run_start[r + 1]), b_end)
            ls = int(L.blk_last_slot[b])
            dl = int(L.slot_deg[ls]) + 1 if ls >= 0 else 0
            rec_ok = dl in PROG_DL
            # cut the run's part into pieces within the header's count limits
            pieces, cur, cnt = [], [], [0] * (nt + 1)
            for pos in range(b, e):
                recs = recs_of[pos]
                add = [0] * nt
                for tp, _, _ in recs or []:
                    add[PROG_TYPES.index(tp)] += 1
                gen_here = recs is None
                over = any(cnt[i] + add[i] > PROG_MAX_COUNT for i in range(nt)) or \
                    (gen_here and cnt[nt] + 1 > PROG_MAX_GEN) or (not gen_here and cnt[nt] > 0)
                if over and cur:
                    pieces.append(cur)
                    cur, cnt = [], [0] * (nt + 1)
                cur.append((pos, recs))
                for i in range(nt):
                    cnt[i] += add[i]
                cnt[nt] += int(gen_here)
            pieces.append(cur)
            for piece in pieces:
                by_type = [[] for _ in range(nt)]
                gen = []
                for pos, recs in piece:
                    if recs is None:
                        gen.append(pos)
                    else:
                        for tp, rr, ff in recs:
                            by_type[PROG_TYPES.index(tp)].append((rr, ff))
                assert gen == list(range(gen[0], gen[0] + len(gen))) if gen else True
                if gen:
                    assert gen[-1] == piece[-1][0], "per-block blocks must end the segment"
                f_off, r_off = len(fidx), len(rows)
                for recs in by_type:
                    for rr, ff in recs:
                        rows.extend(rr)
                        fidx.extend(ff)
                    nrec += len(recs)
                n = [len(rr) for rr in by_type]
                gen_b0 = gen[0] if gen else piece[-1][0] + 1
                dim = int(L.slot_dim[ls]) if ls >= 0 else -1
                w_hdrs.append([ls, f_off, r_off, n[0] | n[1] << 16,
                               n[2] | (dl if rec_ok else 0) << 16 | len(gen) << 24, gen_b0, 0,
                               dim])
                bv = np.zeros(PROG_BASIS)
                if rec_ok:
                    bv[:dl] = L.slot_nodes[ls, :dl]
                    bv[4:4 + dl] = L.slot_w[ls, :dl]
                basis.append(bv)
            b = e
        # groups of PROG_GROUP segments with one record each (DL = 2, one row pair, nothing on
        # the per-block path): flag 1 in word 6 of the group's first header
        for i0 in range(0, len(w_hdrs) - PROG_GROUP + 1, PROG_GROUP):
            if all(h[3] == 1 and (h[4] & 0xffff) == 0 and (h[4] >> 16) == 2
                   for h in w_hdrs[i0:i0 + PROG_GROUP]):
                w_hdrs[i0][6] = 1
        hdrs.extend(w_hdrs)
        seg_off.append(len(hdrs))
    assert len(fidx) < 2**31 and len(rows) < 2**31
    return EvalProgram(seg_off=np.array(seg_off, dtype=np.int32),
                       seg=np.array(hdrs, dtype=np.int64).astype(np.uint32).view(np.int32).reshape(-1)
                       if hdrs else np.zeros(PROG_HDR, dtype=np.int32),
                       basis=np.array(basis).reshape(-1) if basis else np.zeros(PROG_BASIS),
                       rows=np.array(rows or [0], dtype=np.int32),
                       f_index=np.array(fidx or [0], dtype=np.int64), n_records=nrec, ones_row=ones)


def program_values(prog: EvalProgram, F: np.ndarray) -> np.ndarray:
    """The program's F stream: the device F (d_out = 1) gathered in record order, zeros at -1."""
    Ff = np.append(np.asarray(F, dtype=np.float64).reshape(-1), 0.0)
    return Ff[prog.f_index]


def program_rows(prog: EvalProgram, row_node, row_scale):
    """Row map of the program's F stream for sbi_pack_rows_f64: the F row map gathered in record
    order, node -1 (a zero) at padding."""
    fi = prog.f_index
    pad = fi < 0
    node = np.where(pad, -1, np.asarray(row_node)[np.where(pad, 0, fi)]).astype(np.int32)
    scale = np.where(pad, 0.0, np.asarray(row_scale, dtype=np.float64)[np.where(pad, 0, fi)])
    return node, scale

r"""Hierarchical (Newton-surplus) form of a Smolyak interpolant on nested nodes (host, FP64).

For nested node sequences and a downward-closed set Lambda (Lambda_{k,t} is downward closed),
the combination-technique interpolant of the reference
    I^Lambda[f] = sum_nu zeta_nu I^nu[f]            (interpolation.py:264-304)
equals the sum of hierarchical increments
    I^Lambda[f](x) = sum_{nu in Lambda} s_nu prod_{j: nu_j > 0} h^{(j)}_{nu_j}(x_j),
    h^{(j)}_k(x) = prod_{r < k} (x - xi_r) / (xi_k - xi_r)   (Newton basis, h_k(xi_k) = 1),
with the surpluses s_nu obtained from the grid values f(xi_nu) by 1D hierarchisation along
every dimension.  Same polynomial, |Lambda| terms instead of K_raw, and no cancelling zeta
coefficients, which is what lets the mixed-precision kernel meet a 1e-6 relative error.

The node sequence of every dimension must be nested: the 1D nodes of degree k are the first
k + 1 nodes of the sequence (Leja).  Gauss-Hermite (non-nested) is rejected.
"""

import numpy as np

from . import indices


def newton_basis_consts(xi):
    """c_k = 1 / prod_{r<k} (xi_k - xi_r) so that h_k(x) = c_k prod_{r<k} (x - xi_r)."""
    xi = np.asarray(xi, dtype=np.float64)
    c = np.ones(len(xi))
    for k in range(1, len(xi)):
        p = 1.0
        for r in range(k):
            p *= xi[k] - xi[r]
        c[k] = 1.0 / p
    return c


def newton_basis(x, xi, c, k):
    """h_k(x) for an array x."""
    out = np.full(np.shape(x), c[k], dtype=np.float64)
    for r in range(k):
        out = out * (x - xi[r])
    return out


class HierSteps:
    """Hierarchisation as a program: steps in order; step k updates the rows
    tgt[step_off[k]:step_off[k+1]], target i from the rows src[src_off[i]:src_off[i+1]] with
    coefficients coef[...] in that order: s[tgt] = (...((s[tgt] - c0 s[src0]) - c1 s[src1]) ...).
    The layout of sbi_hierarchize_f64's arguments."""

    def __init__(self, step_off, tgt, src_off, src, coef):
        self.step_off, self.tgt, self.src_off, self.src, self.coef = step_off, tgt, src_off, src, coef

    @property
    def n_steps(self):
        return len(self.step_off) - 1


def hierarchy_steps(nus, node_gen):
    """1D hierarchisation along every dimension (ascending), level by level (ascending), of the
    multi-indices `nus` (sparse tuples, downward closed, rows in this order): the target nu with
    nu_j = lv takes the rows of nu with nu_j replaced by l = 0 .. lv - 1 (already final along j),
    with coefficients h_l(xi_lv) of dim j's Newton basis."""
    index = {nu: i for i, nu in enumerate(nus)}
    maxdeg = {}
    for nu in nus:
        for j, v in nu:
            maxdeg[j] = max(maxdeg.get(j, 0), v)
    by_dim = {}
    for nu in nus:
        for pos, (j, v) in enumerate(nu):
            by_dim.setdefault(j, {}).setdefault(v, []).append((nu, pos))
    step_off, tgt, src_off, src, coef = [0], [], [0], [], []
    for j in sorted(by_dim):
        xi = np.asarray(node_gen[j](maxdeg[j]), dtype=np.float64)
        c = newton_basis_consts(xi)
        H = np.zeros((len(xi), len(xi)))  # H[i, l] = h_l(xi_i) for l < i
        for i in range(len(xi)):
            for l in range(i):
                H[i, l] = newton_basis(xi[i], xi, c, l)
        for lv in sorted(by_dim[j]):
            for nu, pos in by_dim[j][lv]:
                tgt.append(index[nu])
                head, tail = nu[:pos], nu[pos + 1:]
                for l in range(lv):
                    key = head + tail if l == 0 else head + ((j, l),) + tail
                    src.append(index[key])
                    coef.append(H[lv, l])
                src_off.append(len(src))
            step_off.append(len(tgt))
    i32 = np.int32
    return HierSteps(np.array(step_off, dtype=i32), np.array(tgt, dtype=i32),
                     np.array(src_off, dtype=i32), np.array(src, dtype=i32),
                     np.array(coef, dtype=np.float64))


def apply_steps(steps, values):
    """Host form of sbi_hierarchize_f64: the program applied to a copy of values (rows x d_out)."""
    s = np.array(values, dtype=np.float64, copy=True)
    for k in range(steps.n_steps):
        for i in range(steps.step_off[k], steps.step_off[k + 1]):
            t = steps.tgt[i]
            v = s[t].copy()
            for q in range(steps.src_off[i], steps.src_off[i + 1]):
                v = v - steps.coef[q] * s[steps.src[q]]
            s[t] = v
    return s


def hierarchize(nus, values, node_gen):
    """Surpluses for the multi-indices `nus` (sparse tuples, downward closed) from grid values.

    values: (|Lambda|, d_out) f at the grid nodes, row order of `nus`."""
    return apply_steps(hierarchy_steps(nus, node_gen), values)


class CapacityError(ValueError):
    """The multi-index set exceeds a size limit of the Newton-form layout."""


class NewtonForm:
    """Host description of the Newton form: multi-indices, surpluses and per-dim node tables."""

    def __init__(self, k, t, node_gen, d_out):
        if not node_gen.is_nested:
            raise ValueError("the Newton (hierarchical) form needs nested nodes (Leja)")
        self.nus = indices.indexset(k, t)
        self.k = np.asarray(k, dtype=np.float64)
        self.d_in = len(k)
        self.d_out = d_out
        self.node_gen = node_gen
        maxdeg = np.zeros(self.d_in, dtype=np.int64)
        for nu in self.nus:
            for j, v in nu:
                maxdeg[j] = max(maxdeg[j], v)
        self.maxdeg = maxdeg

    def grid_point(self, nu, zero):
        x = zero.copy()
        for j, v in nu:
            x[j] = self.node_gen[j](v)[v]
        return x

    def set_values(self, values):
        """values[i] = f at the grid node of nus[i] (rows of d_out)."""
        self.s = hierarchize(self.nus, values, self.node_gen)

    def evaluate(self, x):
        """FP64 numpy evaluation (reference for the kernels' layout tests)."""
        x = np.atleast_2d(np.asarray(x, dtype=np.float64))
        P = x.shape[0]
        cache = {}
        y = np.zeros((P, self.d_out))
        for i, nu in enumerate(self.nus):
            w = np.ones(P)
            for j, v in nu:
                key = (j, v)
                h = cache.get(key)
                if h is None:
                    xi = np.asarray(self.node_gen[j](self.maxdeg[j]), dtype=np.float64)
                    h = newton_basis(x[:, j], xi, newton_basis_consts(xi), v)
                    cache[key] = h
                w = w * h
            y += w[:, None] * self.s[i][None, :]
        return y


N_WAVES = 16  # waves per point tile (kWaves in newton.hip)
# Wave split of the streaming kernels: a group costs its nodes plus GROUP_COST node equivalents
# (latency of its dependent on-chip reads), a dim DIM_COST; tuned on the d_in = 1000 workload
# (balancing by nodes alone left the last wave with 140 one-node groups and twice the run time).
GROUP_COST = 20.0
DIM_COST = 2.0


class NewtonLayout:
    """Device layout of the Newton form (see include/smolyax_amd.h, sbi_newton_desc)."""


def build_newton_layout(nf: NewtonForm, n_waves: int = N_WAVES) -> NewtonLayout:
    L = NewtonLayout()
    nus = nf.nus
    root = nus.index(())
    L.s_root = np.asarray(nf.s[root], dtype=np.float64).reshape(-1)
    parent_of = {nu: nu[:-1] for nu in nus if nu}
    internal = sorted({p for p in parent_of.values() if p}, key=lambda v: (len(v), v))
    row = {(): 0}
    for i, v in enumerate(internal):
        row[v] = i + 1
    L.n_internal = len(internal)
    L.int_parent = np.array([row[v[:-1]] for v in internal] or [0], dtype=np.int32)
    L.int_dim = np.array([v[-1][0] for v in internal] or [0], dtype=np.int32)
    L.int_deg = np.array([v[-1][1] for v in internal] or [0], dtype=np.int32)
    depth = [len(v) for v in internal]
    levels = sorted(set(depth))
    L.level_off = np.array([0] + [int(np.sum(np.array(depth) <= lv)) for lv in levels], dtype=np.int32)
    L.n_levels = len(levels)

    idx = {nu: i for i, nu in enumerate(nus)}
    nodes = sorted((nu for nu in nus if nu), key=lambda v: (v[-1][0], v[-1][1], row[v[:-1]]))
    L.n_nodes = len(nodes)
    L.node_q = np.array([row[v[:-1]] for v in nodes] or [0], dtype=np.int32)
    s = np.array([nf.s[idx[v]] for v in nodes]).reshape(len(nodes), -1) if nodes else np.zeros((1, 1))
    # streaming kernels (d_out == 1) read column 0; the GEMM paths use nf.s directly
    L.node_s64 = np.ascontiguousarray(s[:, 0], dtype=np.float64)
    L.node_s32 = L.node_s64.astype(np.float32)
    # groups by last factor (j, k), in node order
    grp = []
    i = 0
    while i < len(nodes):
        key = nodes[i][-1]
        j0 = i
        while i < len(nodes) and nodes[i][-1] == key:
            i += 1
        grp.append((j0, i, key[0], key[1]))
    L.grp = np.array(grp or [(0, 0, 0, 1)], dtype=np.int32).reshape(-1)
    L.n_groups = len(grp)
    # Newton constants per dim
    kmax1 = int(nf.maxdeg.max()) + 1 if len(nf.maxdeg) else 1
    L.kmax1 = kmax1
    L.dim_xi = np.zeros((nf.d_in, kmax1))
    L.dim_c = np.zeros((nf.d_in, kmax1))
    cache = {}
    for j in range(nf.d_in):
        g = nf.node_gen[j]
        key = (g._table_key(), kmax1)
        v = cache.get(key)
        if v is None:
            xi = np.asarray(g(kmax1 - 1), dtype=np.float64)
            v = (xi, newton_basis_consts(xi))
            cache[key] = v
        L.dim_xi[j], L.dim_c[j] = v
    # cost-balanced split of the groups over the waves, cut only at dim boundaries (a wave owns
    # all groups of its dims, so the incremental Newton product starts at k = 1).  A group costs
    # its nodes plus GROUP_COST (its chain of dependent on-chip reads), a new dim DIM_COST.
    new_dim = [g == 0 or grp[g - 1][2] != grp[g][2] for g in range(len(grp))]
    cost = np.array([(b - a) + GROUP_COST + (DIM_COST if nd else 0.0)
                     for (a, b, _, _), nd in zip(grp, new_dim)]) if grp else np.zeros(0)
    dim_start = [g for g in range(len(grp)) if g == 0 or grp[g - 1][2] != grp[g][2]] + [len(grp)]
    cum = np.concatenate([[0.0], np.cumsum(cost)])
    cuts = [0]
    for w in range(1, n_waves):
        target = cum[-1] * w / n_waves
        pos = int(np.searchsorted(cum[dim_start], target))
        pos = min(max(pos, 0), len(dim_start) - 1)
        cuts.append(dim_start[pos])
    cuts.append(len(grp))
    cuts = np.maximum.accumulate(np.clip(np.array(cuts), 0, len(grp)))
    L.wave_grp = cuts.astype(np.int32)
    nxt = np.full(max(len(grp), 1), -1, dtype=np.int32)
    for w in range(n_waves):
        a, b = int(cuts[w]), int(cuts[w + 1])
        nd = -1
        for g in range(b - 1, a - 1, -1):
            nxt[g] = nd
            if g == a or grp[g - 1][2] != grp[g][2]:
                nd = grp[g][2]
    L.grp_next_dim = nxt

    # staged kernel: groups padded to even counts; pairs {c_deg * s (FP32 bits), parent row
    # byte offset}; 16-byte group records
    s_pad, q_pad, rec = [], [], []
    for a, b, j, k in grp:
        n0 = len(s_pad)
        s_pad.extend((L.node_s64[a:b] * L.dim_c[j, k]).tolist())
        q_pad.extend((L.node_q[a:b].astype(np.int64) * 256).tolist())
        if len(s_pad) % 2:
            s_pad.append(0.0)
            q_pad.append(0)
        rec.append(((n0 & 0xFFFF) | ((len(s_pad) & 0xFFFF) << 16), j, L.dim_xi[j, k - 1]))
    L.n_nodes_pad = len(s_pad)
    # the staged kernel needs node ranges below 65536 and c_deg * s inside FP32's range (c_deg
    # grows like 2^deg on Leja nodes); otherwise its tables are left out and the plan runs the
    # unstaged kernel, which keeps c_deg in the FP64 factor
    folded = np.asarray(s_pad, dtype=np.float64)
    L.node_sq = None
    if L.n_nodes_pad < 65536 and np.all(np.abs(folded) < 1e37):
        sq = np.zeros((max(L.n_nodes_pad, 2), 2), dtype=np.int32)
        sq[: L.n_nodes_pad, 0] = folded.astype(np.float32).view(np.int32)
        sq[: L.n_nodes_pad, 1] = np.array(q_pad, dtype=np.int32)
        L.node_sq = sq.reshape(-1)
    recdt = np.dtype([("n", "<u4"), ("j", "<i4"), ("a", "<f8")])
    L.grp_rec = np.array(rec or [(0, 0, 0.0)], dtype=recdt)
    L.grp_j = np.array([r[1] for r in rec] or [0], dtype=np.int32)
    # h table of the internal nodes' factors: per dim rows k = 1..kmax
    hs_k = {}
    for q in range(L.n_internal):
        j, k = int(L.int_dim[q]), int(L.int_deg[q])
        hs_k[j] = max(hs_k.get(j, 0), k)
    hs_dims = sorted(hs_k)
    rows, r = {}, 0
    for j in hs_dims:
        rows[j] = r
        r += hs_k[j]
    L.n_hs_dims, L.n_hs_rows = len(hs_dims), r
    L.hs_dim = np.array(hs_dims or [0], dtype=np.int32)
    L.hs_kmax = np.array([hs_k[j] for j in hs_dims] or [0], dtype=np.int32)
    L.hs_row = np.array([rows[j] for j in hs_dims] or [0], dtype=np.int32)
    if L.n_internal + 1 < 65536 and r < 65536:
        irec = [int(L.int_parent[q]) | ((rows[int(L.int_dim[q])] + int(L.int_deg[q]) - 1) << 16)
                for q in range(L.n_internal)]
    else:  # 16-bit record fields overflow: no staged kernel (its tables are left out)
        irec, L.node_sq = [], None
    L.int_rec = np.array(irec or [0], dtype=np.int32)
    build_cost_order_tables(nf, L, rows)
    build_basis_tables(nf, L)
    # basis rows for the GEMM paths: the heaviest rows first (FP64 head of the mixed contraction)
    L.basis_desc_nus = L.basis_desc
    set_head(nf, L, HEAD_ROWS)
    return L


def set_head(nf: NewtonForm, L: NewtonLayout, n_head: int):
    """Put the n_head heaviest rows first in the GEMM paths' basis row order (head_order) and
    permute the basis row descriptors to that order."""
    L.basis_order, L.n_head, L.head_fac = head_order(nf, L, n_head)
    desc = L.basis_desc_nus.view(np.uint16).reshape(L.basis_n, 8)
    L.basis_desc = np.ascontiguousarray(desc[L.basis_order]).reshape(-1).view(np.uint8)


LINE_DIMS = 16    # dims per 128-byte line of coordinates


def build_cost_order_tables(nf: NewtonForm, L: NewtonLayout, hs_rows: dict) -> None:
    """Cost-ordered product table and dense surplus vectors (matrix-core mixed kernel, d_out = 1).

    The internal nodes (and the root) are numbered by increasing cost sum_j k_j nu_j: the parents
    of a group (j, k) are the internal nodes q with every dim below j and cost(q) + k k_j within
    the threshold, so for dims past every internal node's dims they are exactly the first rows of
    that order, and for the others a short prefix with gaps.  Each group keeps a dense vector of
    c_k * s (FP32) over the rows 0 .. (its last parent), zeros in the gaps: t = sum_q S_g[q] W[q]
    needs no parent indices.

    Sets: rw_n (rows of W), rint_rec [n_internal] (row | parent row << 16, depth order),
    rint_h [n_internal] (h-table row), and the matrix-core tables (build_mfma_tables); mt_a is
    None when a limit is exceeded (then the staged kernel serves the plan).
    """
    nus = nf.nus
    k = nf.k
    cost = lambda nu: float(sum(k[j] * v for j, v in nu))
    parent_of = {nu: nu[:-1] for nu in nus if nu}
    internal = sorted({p for p in parent_of.values() if p})
    order = sorted([()] + internal, key=lambda v: (cost(v), len(v), v))
    rrow = {v: i for i, v in enumerate(order)}
    L.rw_n = len(order)
    lvl = sorted(internal, key=lambda v: (len(v), rrow[v]))
    L.rint_rec = np.array([rrow[v] | (rrow[v[:-1]] << 16) for v in lvl] or [0], dtype=np.int64)
    L.rint_rec = L.rint_rec.astype(np.uint32).view(np.int32)
    L.rint_h = np.array([hs_rows[v[-1][0]] + v[-1][1] - 1 for v in lvl] or [0], dtype=np.int32)
    g4 = L.grp.reshape(-1, 4)
    node_by_grp = sorted((nu for nu in nus if nu), key=lambda v: (v[-1][0], v[-1][1]))
    idx = {nu: i for i, nu in enumerate(nus)}
    s0 = np.asarray(nf.s, dtype=np.float64).reshape(len(nus), -1)[:, 0]
    vecs = []
    for g in range(L.n_groups):
        n0, n1, j, kk = (int(v) for v in g4[g])
        members = node_by_grp[n0:n1]
        vec = np.zeros(max(rrow[v[:-1]] for v in members) + 1)
        for v in members:
            vec[rrow[v[:-1]]] = s0[idx[v]] * L.dim_c[j, kk]
        vecs.append(vec)
    L.mt_a = None
    if not vecs or L.rw_n >= 65536 or not all(np.all(np.abs(v) < 1e37) for v in vecs):
        return
    build_mfma_tables(L, vecs)


MFMA_WAVES = 4       # wave groups per point tile of the matrix-core mixed kernel (kMGroups)
MFMA_ITEM_STEPS = 16  # k steps per work item at most (longer tiles are split over items)
MFMA_ITEM_COST = 12.0  # per-item work besides its k steps (coordinates, Newton factors, sums)
# Relative rates of the kernel's wave groups 0..3 on equal work (in-kernel cycle counts at
# d_in = 1000: the older waves are issued first, so the groups of higher waves finish later);
# the work items are balanced on these rates.
MFMA_GROUP_SPEED = (1.0, 0.93, 0.68, 0.52)


def build_mfma_tables(L: NewtonLayout, vecs, n_waves: int = MFMA_WAVES) -> None:
    """Tables of the matrix-core mixed kernel (d_out = 1): the groups' dense surplus vectors
    (cost-ordered rows, see build_cost_order_tables) contracted with the product table W on the
    FP32 matrix cores, 16 groups x 4 rows x 16 points per v_mfma_f32_16x16x4.

    Groups are packed 16 to a tile: those whose vectors are prefixes (every dim past the internal
    nodes' dims) in their order (consecutive dims), the others sorted by length; dummy groups
    (dim 0, degree 1, zero surpluses) fill the last tile of each kind.  A tile's A operand is
    stored step by step: for rows 4s .. 4s + 3, 64 floats in MFMA lane order (lane l holds group
    l % 16, row 4s + l / 16), over as many steps as its longest vector needs.  A work item is a
    range of at most MFMA_ITEM_STEPS steps of one tile (y is linear in the partial sums, so a
    tile's steps may be split over items).

    Sets: mt_n (tiles), mt_grp (16-byte group descriptors [16 * mt_n]: {int32 dim, int32 degree,
    double xi_0 of the dim}), mt_a (FP32 A steps, 64 per step), mt_steps, mi_n (work items),
    mi_rec (int32 [4 * mi_n]: first A step, first step within the tile, step count, tile),
    mwave_item [n_waves + 1] (item ranges per wave group, balanced by steps + MFMA_ITEM_COST),
    n_xi_dims (1 + the largest dim of a group of degree >= 2: the kernel stages those dims' node
    rows on chip).
    """
    g4 = L.grp.reshape(-1, 4)
    ng = L.n_groups
    lens = np.array([len(v) for v in vecs])
    last_int_dim = int(L.int_dim[: L.n_internal].max()) if L.n_internal else -1
    pre = [g for g in range(ng) if g4[g, 2] > last_int_dim]
    other = sorted((g for g in range(ng) if g4[g, 2] <= last_int_dim), key=lambda g: -lens[g])
    tiles = []
    for i in range(0, len(other), 16):
        tiles.append(other[i:i + 16])
    # prefix groups: a new tile also where the dims enter a new 128-byte line of coordinates, so
    # that a tile's coordinate reads never share a line with another tile's
    cur = []
    for g in pre:
        if cur and (len(cur) == 16 or g4[g, 2] // LINE_DIMS != g4[cur[-1], 2] // LINE_DIMS):
            tiles.append(cur)
            cur = []
        cur.append(g)
    if cur:
        tiles.append(cur)
    A, desc, tile_a0, tile_ns = [], [], [], []
    s0 = 0
    for t in tiles:
        ns = int(-(-max(lens[g] for g in t) // 4))
        blk = np.zeros((ns, 4, 16))  # [step, row in step, group]
        for i, g in enumerate(t):
            v = np.zeros(4 * ns)
            v[: len(vecs[g])] = vecs[g]
            blk[:, :, i] = v.reshape(ns, 4)
        A.append(blk.reshape(ns, 64))  # lane l = 16 * (row in step) + group
        tile_a0.append(s0)
        tile_ns.append(ns)
        s0 += ns
        for i in range(16):
            if i < len(t):
                j, kk = int(g4[t[i], 2]), int(g4[t[i], 3])
                desc.append((j, kk, L.dim_xi[j, 0]))
            else:
                desc.append((0, 1, 0.0))
    L.mt_n = len(tiles)
    L.mt_a64 = np.concatenate(A).reshape(-1) if A else np.zeros(64)
    L.mt_a = L.mt_a64.astype(np.float32)
    L.mt_steps = s0
    descdt = np.dtype([("j", "<i4"), ("k", "<i4"), ("a", "<f8")])
    L.mt_grp = np.array(desc or [(0, 1, 0.0)] * 16, dtype=descdt)
    items = []
    for ti, (a0, ns) in enumerate(zip(tile_a0, tile_ns)):
        cut = int(-(-ns // MFMA_ITEM_STEPS))
        per = int(-(-ns // cut))
        for c in range(0, ns, per):
            items.append((a0 + c, c, min(per, ns - c), ti))
    cost = [it[2] + MFMA_ITEM_COST for it in items]
    # items that read a common 128-byte line of coordinates (16 consecutive dims) form a cluster
    # and go to one wave group, one after the other, so a line is fetched from memory once per
    # tile and then found in the cache
    dims = np.array([g[0] for g in desc], dtype=np.int64).reshape(-1, 16)
    root = list(range(len(items)))

    def find(a):
        while root[a] != a:
            root[a] = root[root[a]]
            a = root[a]
        return a
    line_item = {}
    for ii, it in enumerate(items):
        for ln in set((dims[it[3]] // LINE_DIMS).tolist()):
            if ln in line_item:
                root[find(ii)] = find(line_item[ln])
            else:
                line_item[ln] = ii
    clusters = {}
    for ii in range(len(items)):
        clusters.setdefault(find(ii), []).append(ii)
    clusters = list(clusters.values())
    ccost = [sum(cost[i] for i in c) for c in clusters]
    load = np.zeros(n_waves)
    speed = np.array(MFMA_GROUP_SPEED[:n_waves]) if n_waves == len(MFMA_GROUP_SPEED) else \
        np.ones(n_waves)
    owner = [[] for _ in range(n_waves)]
    for ci in sorted(range(len(clusters)), key=lambda c: -ccost[c]):
        w = int(np.argmin((load + ccost[ci]) / speed))
        owner[w].append(ci)
        load[w] += ccost[ci]
    L.mwave_item = np.array(np.cumsum([0] + [sum(len(clusters[c]) for c in o) for o in owner]),
                            dtype=np.int32)
    items = [items[i] for w in range(n_waves) for c in sorted(owner[w], key=lambda c: min(clusters[c]))
             for i in clusters[c]]
    L.mi_n = len(items)
    L.mi_rec = np.array(items or [(0, 0, 0, 0)], dtype=np.int32).reshape(-1)
    high = g4[g4[:, 3] >= 2, 2]
    L.n_xi_dims = int(high.max()) + 1 if len(high) else 0


HEAD_ROWS = 64  # rows of the Newton basis that the mixed GEMM path contracts in FP64


def head_order(nf: NewtonForm, L: NewtonLayout, n_head: int = HEAD_ROWS):
    """Row order of the basis matrix on the GEMM paths, and the factors of the head rows.

    The n_head rows with the largest max_x |h_nu(x)| * ||s_nu||_2 (max over each dim's domain, on
    a 513-point grid) come first, the others follow in nus order.  These few rows (the constant and
    the low-degree terms of the first dims) carry most of the result; the mixed contraction keeps
    them in FP64 so that their FP32 rounding does not dominate its error.
    Returns (order into nf.nus, n_head, int32 [16 * n_head] (dim, deg) pairs, dim -1 = none).
    """
    nus = nf.nus
    n_head = min(n_head, len(nus))
    grid = np.linspace(-1.0, 1.0, 513)
    peak = {}

    def hmax(j, k):
        v = peak.get((j, k))
        if v is None:
            xs = np.asarray(nf.node_gen[j].scale(grid), dtype=np.float64)
            h = np.full(grid.shape, L.dim_c[j, k])
            for r in range(k):
                h = h * (xs - L.dim_xi[j, r])
            v = peak[(j, k)] = float(np.max(np.abs(h)))
        return v

    snorm = np.linalg.norm(np.asarray(nf.s, dtype=np.float64).reshape(len(nus), -1), axis=1)
    w = np.array([np.prod([hmax(j, k) for j, k in nu]) if nu else 1.0 for nu in nus]) * snorm
    head = np.argsort(-w, kind="stable")[:n_head]
    is_head = np.zeros(len(nus), dtype=bool)
    is_head[head] = True
    order = np.concatenate([head, np.nonzero(~is_head)[0]]).astype(np.int64)
    fac = np.zeros((max(n_head, 1), 8, 2), dtype=np.int32)
    fac[..., 0] = -1
    for r, i in enumerate(head):
        for q, (j, k) in enumerate(nus[i]):
            fac[r, q] = (j, k)
    return order, n_head, fac.reshape(-1)


JAC_UNIT = 128  # adjoint entries per work unit of the Jacobian's register pass


def build_basis_tables(nf: NewtonForm, L: NewtonLayout) -> None:
    """Descriptors of the basis matrix H^T (rows in nf.nus order) and of the adjoint."""
    tab_dims = [j for j in range(nf.d_in) if nf.maxdeg[j] > 0]
    row0, r = {}, 0
    for j in tab_dims:
        row0[j] = r
        r += int(nf.maxdeg[j])
    if r >= 65535:
        raise CapacityError("factor table too large (at most 65534 rows)")
    L.n_tab_dims, L.n_tab_rows = len(tab_dims), r
    L.tab_dim = np.array(tab_dims or [0], dtype=np.int32)
    L.tab_row = np.array([row0[j] for j in tab_dims] or [0], dtype=np.int32)
    L.tab_kmax = np.array([int(nf.maxdeg[j]) for j in tab_dims] or [0], dtype=np.int32)
    L.basis_n = len(nf.nus)
    desc = np.full((max(L.basis_n, 1), 8), 0xFFFF, dtype=np.uint16)
    per_row = [[] for _ in range(max(r, 1))]
    for i, nu in enumerate(nf.nus):
        if len(nu) > 8:
            raise CapacityError("basis descriptors hold at most 8 factors per node")
        rows = [row0[j] + k - 1 for j, k in nu]
        desc[i, : len(rows)] = rows
        for a, ra in enumerate(rows):
            others = rows[:a] + rows[a + 1:]
            if len(others) > 6:
                raise CapacityError("adjoint entries hold at most 6 other factors")
            per_row[ra].append((i, others))
    L.basis_desc = desc.reshape(-1).view(np.uint8)
    off = [0]
    ent = []
    for ra in range(r):
        for node, others in per_row[ra]:
            e = np.full(8, 0xFFFF, dtype=np.uint16)
            e[0:2] = np.array([node], dtype=np.int32).view(np.uint16)
            e[2: 2 + len(others)] = others
            ent.append(e)
        off.append(len(ent))
    L.vjp_slot_off = np.array(off, dtype=np.int32)
    L.n_vjp_ent = len(ent)
    L.vjp_ent_row = np.repeat(np.arange(max(r, 1), dtype=np.int32),
                              np.diff(np.asarray(off)) if r else [0]).astype(np.int32)
    if not len(L.vjp_ent_row):
        L.vjp_ent_row = np.zeros(1, dtype=np.int32)
    L.row_tdim = np.zeros(max(r, 1), dtype=np.int32)
    for i, j in enumerate(tab_dims):
        L.row_tdim[row0[j]: row0[j] + int(nf.maxdeg[j])] = i
    # entries are stored by factor row, rows by dim: CSR of the entries over input dims
    dim_off = np.zeros(nf.d_in + 1, dtype=np.int64)
    for j in tab_dims:
        dim_off[j + 1] = off[row0[j] + int(nf.maxdeg[j])] - off[row0[j]]
    L.vjp_dim_off = np.cumsum(dim_off).astype(np.int32)
    # work units of the Jacobian's register pass: runs of at most JAC_UNIT entries of one dim;
    # a dim split into several units sums their partial results afterwards (fixed order)
    units, red, n_part = [], [], 0
    for i, j in enumerate(tab_dims):
        e0, e1 = int(L.vjp_dim_off[j]), int(L.vjp_dim_off[j + 1])
        cuts = list(range(e0, e1, JAC_UNIT)) + [e1]
        if len(cuts) <= 2:
            units.append((i, e0, e1, -1))
        else:
            red.append((j, n_part, len(cuts) - 1))
            for a, b in zip(cuts[:-1], cuts[1:]):
                units.append((i, a, b, n_part))
                n_part += 1
    L.n_jac_units, L.n_jac_partials, L.n_jac_red = len(units), n_part, len(red)
    L.jac_unit = np.array(units or [(0, 0, 0, -1)], dtype=np.int32).reshape(-1)
    L.jac_red = np.array(red or [(0, 0, 0)], dtype=np.int32).reshape(-1)
    L.vjp_ent = (np.stack(ent) if ent else np.full((1, 8), 0xFFFF, dtype=np.uint16)).reshape(-1).view(np.uint8)


def eval_newton_layout(L: NewtonLayout, x, fp32=False):
    """Numpy model of the kernel's arithmetic (test helper for the host layout and precision)."""
    x = np.atleast_2d(np.asarray(x, dtype=np.float64))
    P = x.shape[0]
    T = np.float32 if fp32 else np.float64
    W = np.zeros((L.n_internal + 1, P), dtype=T)
    W[0] = 1
    for q in range(L.n_internal):
        j, k = L.int_dim[q], L.int_deg[q]
        h = np.full(P, L.dim_c[j, k])
        for r in range(k):
            h = h * (x[:, j] - L.dim_xi[j, r])
        W[q + 1] = (W[L.int_parent[q]].astype(np.float64) * h).astype(T)
    s = (L.node_s32 if fp32 else L.node_s64)
    acc = np.zeros(P)
    g4 = L.grp.reshape(-1, 4)
    starts = set(int(v) for v in L.wave_grp[:-1])
    jcur, kdone, pr = -1, 0, None
    for g in range(L.n_groups):
        n0, n1, j, k = g4[g]
        # incremental Newton product as in the kernel (restarted at every wave's first group)
        if j != jcur or g in starts:
            jcur, kdone, pr = j, 0, np.ones(P)
        while kdone < k:
            pr = pr * (x[:, j] - L.dim_xi[j, kdone])
            kdone += 1
        h = pr * L.dim_c[j, k]
        t = np.zeros(P, dtype=T)
        for i in range(n0, n1):
            t = (t + s[i] * W[L.node_q[i]]).astype(T)
        acc += h * t.astype(np.float64)
    return acc + L.s_root[0]


def eval_mfma_layout(L: NewtonLayout, x):
    """Numpy model of the matrix-core mixed kernel (test helper): per tile, T = A^T W over its
    steps in FP32, then y += h_k(x_j) T in FP64 with h_k(x) = prod_{r<k} (x - xi_r) (the c_k of
    the Newton basis is folded into the surpluses)."""
    x = np.atleast_2d(np.asarray(x, dtype=np.float64))
    P = x.shape[0]
    h = np.zeros((max(L.n_hs_rows, 1), P), dtype=np.float32)
    for i in range(L.n_hs_dims):
        j, km, r0 = int(L.hs_dim[i]), int(L.hs_kmax[i]), int(L.hs_row[i])
        pr = np.ones(P)
        for kk in range(1, km + 1):
            pr = pr * (x[:, j] - L.dim_xi[j, kk - 1])
            h[r0 + kk - 1] = (pr * L.dim_c[j, kk]).astype(np.float32)
    W = np.zeros((-(-L.rw_n // 4) * 4, P), dtype=np.float32)
    W[0] = 1.0
    rec = L.rint_rec.view(np.uint32)
    for i in range(L.n_internal):
        r, par = int(rec[i] & 0xFFFF), int(rec[i] >> 16)
        W[r] = W[par] * h[int(L.rint_h[i])]
    A = L.mt_a.reshape(-1, 4, 16)
    acc = np.zeros(P)
    for it in range(L.mi_n):
        a0, w0, ns, t = (int(v) for v in L.mi_rec[4 * it: 4 * it + 4])
        T = np.zeros((16, P), dtype=np.float32)
        for s_ in range(ns):
            rows = W[4 * (w0 + s_): 4 * (w0 + s_) + 4]
            T = (T + A[a0 + s_].T.astype(np.float32) @ rows).astype(np.float32)
        for i in range(16):
            dsc = L.mt_grp[16 * t + i]
            j, kk = int(dsc["j"]), int(dsc["k"])
            pr = np.ones(P)
            for r in range(kk):
                pr = pr * (x[:, j] - L.dim_xi[j, r])
            acc += pr * T[i].astype(np.float64)
    return acc + L.s_root[0]

"""Oracle restatement of the reference Smolyak interpolant (FP64, numpy).

Follows src/smolyax/interpolation.py:115-390 (set_f packing into padded groups, __call__,
gradient, integral) and src/smolyax/barycentric.py:34-229 (basis, tensor-product value and
gradient).  jax.vmap over the blocks of a group becomes a leading block axis; jnp.einsum becomes
numpy.einsum.  Test infrastructure only.

Pinned by (tests/test_oracle_smolyak.py):
  * polynomial reproduction I[p] = p for p in P_Lambda (tests/test_smolyak.py:7-26 of the
    reference, Legendre / Hermite tensor polynomials),
  * exact interpolation at grid nodes (one-hot branch),
  * quadrature exactness, integral of P_nu = delta_{nu,0} (tests/test_smolyak.py:29-49),
  * gradient exactness against analytic derivatives (tests/test_smolyak.py:52-76) and central
    finite differences; NaN at grid nodes,
  * sum of zeta = 1 (constant reproduction).

Extended precision: `__call__`, `gradient` and `integral` take dtype=np.longdouble (x87 80-bit,
64-bit significand on x86-64).  The FP64 data of the interpolant (node values, f at the nodes,
quadrature weights) is used as it is; the barycentric weights are recomputed from the FP64 node
values in the extended type (`compute_weights_ext`), so the second-form basis is the Lagrange
basis of those nodes to ~1e-19, and every product and sum runs in the extended type.  The result
is the exact polynomial interpolant of the FP64 data up to ~1e-19 * sum_nu |zeta_nu| |I^nu|,
far below the FP64 rounding (~1e-16 * the same sum) it serves to measure.  Pinned by
tests/test_oracle_extended.py: it agrees with exact rational arithmetic (fractions) on small
sets, and with the FP64 mode to the FP64 rounding the coefficients amplify.
"""

import itertools as it
import string
from collections import defaultdict

import numpy as np

from . import indices
from .nodes import compute_weights

EXT = np.longdouble


def compute_weights_ext(nodes):
    """barycentric.py:13-31 in the extended type, from the FP64 node values."""
    nodes = np.asarray(nodes, dtype=np.float64).astype(EXT)
    diffs = nodes[:, None] - nodes
    diffs = np.where(diffs == 0, EXT(1), diffs)
    return np.prod(EXT(1) / diffs, axis=0)


def _basis_unnormalized(x, xi, w, nu_i):
    """barycentric.py:34-66.  x (..., 1), xi/w (..., m) -> (..., m)."""
    diffs = x - xi
    m = diffs.shape[-1]
    mask_cols = np.arange(m) <= nu_i[..., None]
    zero = diffs == 0
    mask_zero = np.any(zero & mask_cols, axis=-1, keepdims=True)
    with np.errstate(divide="ignore", invalid="ignore"):
        wd = w / diffs
    b = np.where(mask_zero, zero.astype(diffs.dtype), wd)
    return np.where(mask_cols, b, 0.0)


def _basis_gradient_unnormalized(x, xi, w, nu_i):
    """barycentric.py:126-155."""
    diffs = (x - xi) ** 2
    m = diffs.shape[-1]
    mask_cols = np.arange(m) <= nu_i[..., None]
    singular = (diffs == 0) & mask_cols
    with np.errstate(divide="ignore", invalid="ignore"):
        wd = -w / diffs
    wd = np.where(singular, np.nan, wd)
    return np.where(mask_cols, wd, 0.0)


class Interpolant:
    """Restatement of SmolyakBarycentricInterpolator (interpolation.py:15-390)."""

    def __init__(self, *, d_out, node_gen, k, t, f=None, memory_limit=4.0):
        self.d_in = len(k)
        self.d_out = d_out
        self.node_gen = node_gen
        self.is_nested = node_gen.is_nested
        self.k = np.asarray(k, dtype=float)
        self.t = t
        self.memory_limit = memory_limit
        self.offset = 0
        self.groups = {}
        self.n_f_evals = indices.nodeset_cardinality(self.k, t, nested=self.is_nested)
        self.n_f_evals_new = 0
        if f is not None:
            self.set_f(f=f)

    # ---------------------------------------------------------------- set_f (interpolation.py:115-239)
    def set_f(self, *, f, f_evals=None):
        if f_evals is None:
            f_evals = {}
        zero = np.array([g(0)[0] for g in self.node_gen])
        n2nus, n2zetas = indices.non_zero_indices_and_zetas(self.k, self.t)
        self.groups = {}
        for n, nus in n2nus.items():
            if n == 0:
                nu = ()
                fe = f_evals if self.is_nested else f_evals.get(nu, {})
                if nu not in fe:
                    fe[nu] = f(zero.copy())
                    self.n_f_evals_new += 1
                if not self.is_nested:
                    f_evals[nu] = fe
                self.offset = n2zetas[n][0] * np.asarray(fe[nu], dtype=float)
                continue
            nn = len(nus)
            sdims = np.empty((nn, n), dtype=int)
            sdegs = np.empty((nn, n), dtype=int)
            for i, nu in enumerate(nus):
                snu = sorted(nu, key=lambda x: x[1], reverse=True)
                sdims[i], sdegs[i] = zip(*snu)
            tau = tuple(int(v) for v in sdegs.max(axis=0))
            tmax = max(tau)
            nodes = np.zeros((nn, n, tmax + 1))
            weights = np.zeros((nn, n, tmax + 1))
            weights_ext = np.zeros((nn, n, tmax + 1), dtype=EXT)
            for s in range(n):
                grp = defaultdict(list)
                for i in range(nn):
                    grp[(int(sdims[i, s]), int(sdegs[i, s]))].append(i)
                for (dim, deg), idxs in grp.items():
                    pts = self.node_gen[dim](deg)
                    nodes[idxs, s, : deg + 1] = pts
                    weights[idxs, s, : deg + 1] = compute_weights(pts)
                    weights_ext[idxs, s, : deg + 1] = compute_weights_ext(pts)
            F = np.zeros((nn,) + tuple(v + 1 for v in tau) + (self.d_out,))
            for i, nu in enumerate(nus):
                x = zero.copy()
                fe = f_evals if self.is_nested else f_evals.get(nu, {})
                s_i = sdims[i]
                order = np.argsort(s_i)
                for mu in it.product(*[range(v + 1) for v in sdegs[i]]):
                    key = tuple((s_i[q], mu[q]) for q in order if mu[q] > 0)
                    if key not in fe:
                        x[s_i] = [nodes[i, q, mu[q]] for q in range(n)]
                        fe[key] = f(x)
                        self.n_f_evals_new += 1
                    F[i][mu] = fe[key]
                if not self.is_nested:
                    f_evals[nu] = fe
            self.groups[n] = dict(
                F=np.moveaxis(F, -1, 1), nodes=nodes, weights=weights, weights_ext=weights_ext,
                dims=sdims, degs=sdegs,
                zetas=np.asarray(n2zetas[n], dtype=float), tau=tau, nus=list(nus),
            )
        return f_evals

    # ---------------------------------------------------------------- helpers
    def _batches(self, n, per_summand_elems, itemsize=8):
        """Summand batches under memory_limit GiB (interpolation.py:285-287, 327-329)."""
        nn = self.groups[n]["F"].shape[0]
        mem = per_summand_elems * itemsize / 1024**3
        bs = max(1, int(np.floor(self.memory_limit / mem)))
        for s in range(0, nn, bs):
            yield s, min(s + bs, nn)

    @staticmethod
    def _slot(x, g, a, b, s, m, dt):
        """Coordinates, nodes and barycentric weights of slot s of blocks a..b in dtype dt."""
        ext = np.dtype(dt) != np.float64
        xs = x[:, g["dims"][a:b, s]].T[:, :, None].astype(dt)          # (B, P, 1)
        xi = g["nodes"][a:b, s, None, :m].astype(dt)                    # (B, 1, m)
        w = (g["weights_ext"] if ext else g["weights"])[a:b, s, None, :m]
        return xs, xi, w

    def _norm_bases(self, x, g, a, b, dt=np.float64):
        """Per slot: normalised basis (blocks, P, m_s) as in barycentric.py:114-117."""
        out = []
        for s, m in enumerate(g["tau"]):
            m = m + 1
            xs, xi, w = self._slot(x, g, a, b, s, m, dt)
            nu = g["degs"][a:b, s][:, None]                              # (B, 1)
            bb = _basis_unnormalized(xs, xi, w, nu)
            out.append(bb / bb.sum(axis=-1, keepdims=True))
        return out

    def _prepare(self, x):
        x = np.asarray(x, dtype=float)
        if x.shape == (self.d_in,):
            x = x[None, :]
        assert x.shape[1] == self.d_in, f"{x.shape[1]} != {self.d_in}"
        return x

    # ---------------------------------------------------------------- __call__ (interpolation.py:264-304)
    def __call__(self, x, dtype=np.float64):
        assert self.groups or self.d_in == 0 or self.offset is not None
        x = self._prepare(x)
        P = x.shape[0]
        dt = np.dtype(dtype)
        y = np.broadcast_to(np.asarray(self.offset, dtype=dt), (P, self.d_out)).astype(dt).copy()
        for n, g in self.groups.items():
            letters = string.ascii_letters
            fa = letters[:n]
            subs = "zo" + fa + "," + ",".join("zp" + c for c in fa) + "->zpo"
            for a, b in self._batches(n, P * self.d_out * np.prod(g["F"].shape[3:]), dt.itemsize):
                bs = self._norm_bases(x, g, a, b, dt)
                r = np.einsum(subs, g["F"][a:b].astype(dt), *bs, optimize=True)
                y += np.einsum("z,zpo->po", g["zetas"][a:b].astype(dt), r)
        return y

    # ---------------------------------------------------------------- gradient (interpolation.py:306-345)
    def gradient(self, x, dtype=np.float64):
        x = self._prepare(x)
        P = x.shape[0]
        dt = np.dtype(dtype)
        J = np.zeros((P, self.d_out, self.d_in), dtype=dt)
        for n, g in self.groups.items():
            letters = string.ascii_letters
            fa = letters[:n]
            subs = "zo" + fa + "," + ",".join("zpq" + c for c in fa) + "->zpoq"
            for a, b in self._batches(n, J.size * np.prod(g["F"].shape[3:]), dt.itemsize):
                Bs = []
                for s, m in enumerate(g["tau"]):
                    m = m + 1
                    xs, xi, w = self._slot(x, g, a, b, s, m, dt)
                    nu = g["degs"][a:b, s][:, None]
                    bb = _basis_unnormalized(xs, xi, w, nu)
                    bn = bb.sum(axis=-1, keepdims=True)
                    db = _basis_gradient_unnormalized(xs, xi, w, nu)
                    dbn = db.sum(axis=-1, keepdims=True)
                    bh = bb / bn
                    Bi = db / bn - bh * dbn / bn                              # barycentric.py:219
                    B = np.repeat(bh[:, :, None, :], n, axis=2)                  # (B, P, n, m)
                    B[:, :, s, :] = Bi
                    Bs.append(B)
                r = np.einsum(subs, g["F"][a:b].astype(dt), *Bs, optimize=True)  # (B, P, o, n)
                r = r * g["zetas"][a:b, None, None, None].astype(dt)
                for q in range(b - a):
                    J[:, :, g["dims"][a + q]] += r[q]
        return J

    # ---------------------------------------------------------------- integral (interpolation.py:347-390)
    def integral(self, dtype=np.float64):
        dt = np.dtype(dtype)
        Q = np.broadcast_to(np.asarray(self.offset, dtype=dt), (self.d_out,)).astype(dt).copy()
        for n, g in self.groups.items():
            nn = g["F"].shape[0]
            wl = [np.zeros((nn, v + 1), dtype=dt) for v in g["tau"]]
            for s in range(n):
                grp = defaultdict(list)
                for i in range(nn):
                    grp[(int(g["dims"][i, s]), int(g["degs"][i, s]))].append(i)
                for (dim, deg), idxs in grp.items():
                    q = self.node_gen[dim].get_quadrature_weights(deg)
                    wl[s][idxs, : len(q)] = q
            fa = string.ascii_lowercase[:n]
            o = string.ascii_lowercase[n]
            subs = "z," + ",".join("z" + c for c in fa) + ",z" + o + fa + "->" + o
            Q += np.einsum(subs, g["zetas"].astype(dt), *wl, g["F"].astype(dt), optimize=True)
        return Q

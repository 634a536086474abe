r"""Anisotropic multi-index sets $\Lambda_{k,t} = \{\nu : \sum_j k_j \nu_j < t\}$ (src/smolyax/indices.py).

The work runs in native host code (libsmolyax_amd.so), which restates the reference's numba kernels
with the same traversal order and IEEE operation sequence, so the sets, cardinalities, Smolyak
coefficients and thresholds are bit-identical.  Return types follow the reference.
"""

from collections import defaultdict
from typing import Sequence

import numpy as np

from . import _native as N


def _k(k):
    k = np.ascontiguousarray(np.asarray(k, dtype=np.float64))
    return k, len(k)


def indexset(k: Sequence[float], t: float) -> list:
    """indices.py:20-69: sparse multi-indices ((j, nu_j), ...) in the reference's DFS order."""
    k, d = _k(k)
    n_nu, n_ent = N.c_i64(), N.c_i64()
    N.check(N.fn["sbi_indexset"](N.f64_ptr(k), d, float(t), n_nu, n_ent, None, None, None), "indexset")
    off = np.empty(n_nu.value + 1, dtype=np.int64)
    dims = np.empty(max(n_ent.value, 1), dtype=np.int32)
    degs = np.empty(max(n_ent.value, 1), dtype=np.int32)
    N.check(N.fn["sbi_indexset"](N.f64_ptr(k), d, float(t), n_nu, n_ent, N.i64_ptr(off),
                                 N.i32_ptr(dims), N.i32_ptr(degs)), "indexset")
    dl, gl = dims.tolist(), degs.tolist()
    return [tuple(zip(dl[off[i]:off[i + 1]], gl[off[i]:off[i + 1]])) for i in range(n_nu.value)]


def indexset_cardinality(k: Sequence[float], t: float) -> int:
    """indices.py:72-118."""
    k, d = _k(k)
    return int(N.fn["sbi_indexset_cardinality"](N.f64_ptr(k), d, float(t)))


def smolyak_coefficient(k: Sequence[float], d: int, rem_t: float, parity: int) -> int:
    """indices.py:121-168."""
    k, _ = _k(k)
    return int(N.fn["sbi_smolyak_coefficient"](N.f64_ptr(k), int(d), float(rem_t), int(parity)))


def nonzero_blocks(k: Sequence[float], t: float):
    """Flat form of non_zero_indices_and_zetas: (offsets, dims, degs, zetas) in discovery order."""
    k, d = _k(k)
    n_nu, n_ent = N.c_i64(), N.c_i64()
    f = N.fn["sbi_nonzero_indices_and_zetas"]
    N.check(f(N.f64_ptr(k), d, float(t), n_nu, n_ent, None, None, None, None), "nonzero_indices")
    off = np.empty(n_nu.value + 1, dtype=np.int64)
    dims = np.empty(max(n_ent.value, 1), dtype=np.int32)
    degs = np.empty(max(n_ent.value, 1), dtype=np.int32)
    zet = np.empty(max(n_nu.value, 1), dtype=np.int64)
    N.check(f(N.f64_ptr(k), d, float(t), n_nu, n_ent, N.i64_ptr(off), N.i32_ptr(dims),
              N.i32_ptr(degs), N.i64_ptr(zet)), "nonzero_indices")
    return off, dims[: n_ent.value], degs[: n_ent.value], zet[: n_nu.value]


def non_zero_indices_and_zetas(k: Sequence[float], t: float):
    """indices.py:214-259: (n -> list of nu, n -> list of zeta), nu with zeta != 0 only."""
    off, dims, degs, zet = nonzero_blocks(k, t)
    n2nus, n2zetas = defaultdict(list), defaultdict(list)
    dl, gl = dims.tolist(), degs.tolist()
    for i in range(len(zet)):
        nu = tuple(zip(dl[off[i]:off[i + 1]], gl[off[i]:off[i + 1]]))
        n2nus[len(nu)].append(nu)
        n2zetas[len(nu)].append(int(zet[i]))
    return n2nus, n2zetas


def nodeset_cardinality(k: Sequence[float], t: float, nested: bool = False) -> int:
    """indices.py:262-289."""
    k, d = _k(k)
    return int(N.fn["sbi_nodeset_cardinality"](N.f64_ptr(k), d, float(t), int(bool(nested))))


def find_approximate_threshold(k: Sequence[float], m: int, nested: bool, max_iter: int = 32,
                               accuracy: float = 0.001) -> float:
    """indices.py:292-355."""
    assert m > 0
    if m == 1:
        return 1
    k, d = _k(k)
    return float(N.fn["sbi_find_approximate_threshold"](N.f64_ptr(k), d, int(m), int(bool(nested)),
                                                        int(max_iter), float(accuracy)))

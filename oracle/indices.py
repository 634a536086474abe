"""Oracle restatement of the reference's multi-index machinery (src/smolyax/indices.py).

Every function keeps the reference's traversal order and its exact sequence of IEEE double
operations (budgets are subtracted / accumulated in the same order), so results are bit-identical
to the numba originals, which run without fast-math.

Pinned by (tests/test_oracle_indices.py):
  * brute force over the dense box (restating tests/indices_dense_reference_impl.py:4-29),
  * the reference's known cardinalities d=10000, k_j=log2(2+j): t=5.3 -> 98, t=7.78 -> 1000,
    t=10.317 -> 10003 (tests/test_indices_runtime.py:10-12),
  * sum of Smolyak coefficients = 1 and zeta vs the inclusion-exclusion definition.
"""

from collections import defaultdict

import numpy as np


def indexset(k, t):
    """Sparse multi-indices of Lambda_{k,t}; DFS order of indices.py:20-69."""
    d = len(k)
    stack = [(0, t, ())]
    result = []
    while stack:
        i, remaining_t, nu_head = stack.pop()
        if i >= d or k[i] >= remaining_t:
            result.append(nu_head)
            continue
        stack.append((i + 1, remaining_t, nu_head))
        j = 1
        k_i = k[i]
        while j * k_i < remaining_t:
            stack.append((i + 1, remaining_t - j * k_i, nu_head + ((i, j),)))
            j += 1
    return result


def indexset_cardinality(k, t):
    """|Lambda_{k,t}| by the used-budget DFS of indices.py:72-118."""
    k = np.asarray(k, dtype=np.float64)
    d = len(k)
    stack = [(0, 0.0)]
    count = 0
    while stack:
        dim_i, used_t = stack.pop()
        if dim_i >= d:
            count += 1
            continue
        remaining_t = t - used_t
        if dim_i + 1 < d and k[dim_i + 1] < remaining_t:
            stack.append((dim_i + 1, used_t))
        else:
            count += 1
        j = 1
        while used_t + j * k[dim_i] < t:
            stack.append((dim_i + 1, used_t + j * k[dim_i]))
            j += 1
    return count


def smolyak_coefficient(k, d, rem_t, parity):
    """zeta_nu from the remaining budget by the parity DFS of indices.py:121-168."""
    total = 0
    stack = [(0, rem_t, parity)]
    while stack:
        i, rt, p = stack.pop()
        if i >= d:
            total += 1 - (p << 1)
            continue
        if i + 1 < d and k[i + 1] < rt:
            stack.append((i + 1, rt, p))
        else:
            total += 1 - (p << 1)
        cost = k[i]
        if cost < rt:
            stack.append((i + 1, rt - cost, p ^ 1))
    return total


def _nodeset_cardinality_non_nested(k, t):
    """Sum of prod(nu_j+1) over nu with zeta != 0; indices.py:175-211."""
    k = np.asarray(k, dtype=np.float64)
    d = len(k)
    stack = [(0, t, 0, 1)]
    total = 0
    while stack:
        dim_i, rem_t, parity, prod_n = stack.pop()
        if dim_i >= d or not (dim_i + 1 < d and k[dim_i + 1] < rem_t):
            if smolyak_coefficient(k, d, rem_t, parity) != 0:
                total += prod_n
        if dim_i < d:
            if dim_i + 1 < d and k[dim_i + 1] < rem_t:
                stack.append((dim_i + 1, rem_t, parity, prod_n))
            cost = k[dim_i]
            j = 1
            while cost * j < rem_t:
                stack.append((dim_i + 1, rem_t - cost * j, parity ^ (j & 1), prod_n * (j + 1)))
                j += 1
    return total


def non_zero_indices_and_zetas(k, t):
    """Multi-indices with zeta != 0 and their zetas, binned by #active dims; indices.py:214-259."""
    k = np.asarray(k, dtype=np.float64)
    d = len(k)
    n2nus, n2zetas = defaultdict(list), defaultdict(list)
    stack = [(0, t, ())]
    while stack:
        i, rem_t, nu = stack.pop()
        if i >= d or not (i + 1 < d and k[i + 1] < rem_t):
            zeta = smolyak_coefficient(k, d, rem_t, 0)
            if zeta != 0:
                n = len(nu)
                n2nus[n].append(nu)
                n2zetas[n].append(zeta)
        if i < d:
            if i + 1 < d and k[i + 1] < rem_t:
                stack.append((i + 1, rem_t, nu))
            j = 1
            while j * k[i] < rem_t:
                stack.append((i + 1, rem_t - j * k[i], nu + ((i, j),)))
                j += 1
    return n2nus, n2zetas


def nodeset_cardinality(k, t, nested=False):
    """indices.py:262-289."""
    if nested:
        return indexset_cardinality(k, t)
    return _nodeset_cardinality_non_nested(k, t)


def find_approximate_threshold(k, m, nested, max_iter=32, accuracy=0.001):
    """Bracket (x1.2 steps) then bisect; indices.py:292-355."""
    assert m > 0
    if m == 1:
        return 1
    l_interval = [1.0, 2.0]
    while nodeset_cardinality(k, l_interval[0], nested) > m:
        l_interval = [l_interval[0] / 1.2, l_interval[0]]
    while nodeset_cardinality(k, l_interval[1], nested) < m:
        l_interval = [l_interval[1], l_interval[1] * 1.2]

    def midpoint(interval):
        return interval[0] + (interval[1] - interval[0]) / 2.0

    t_cand = midpoint(l_interval)
    m_cand = nodeset_cardinality(k, t_cand, nested)
    for _ in range(max_iter):
        if m_cand > m:
            l_interval = [l_interval[0], t_cand]
        else:
            l_interval = [t_cand, l_interval[1]]
        t_cand = midpoint(l_interval)
        m_cand = nodeset_cardinality(k, t_cand, nested)
        if np.abs(m_cand - m) / m < accuracy:
            break
    return t_cand


# ---- brute force (tiny inputs only), restating tests/indices_dense_reference_impl.py:4-29 ----

def dense_indexset(k, t, idx=()):
    if len(k) == 0:
        return [idx]
    r = []
    j = 0
    while j * k[0] < t:
        r += dense_indexset(k[1:], t - j * k[0], idx + (j,))
        j += 1
    return r


def _unitball(nu, k, t, e=()):
    if len(k) == 0:
        return [e]
    r = _unitball(nu[1:], k[1:], t - nu[0] * k[0], e + (0,))
    if np.dot(nu, k) + k[0] < t:
        r += _unitball(nu[1:], k[1:], t - (nu[0] + 1) * k[0], e + (1,))
    return r


def dense_smolyak_coefficient(k, t, nu):
    return int(np.sum([(-1) ** (np.sum(e)) for e in _unitball(tuple(nu), tuple(k), t)]))


def sparse_to_dense(nu, d):
    out = [0] * d
    for j, v in nu:
        out[j] = v
    return tuple(out)


def dense_to_sparse(nu):
    return tuple((j, v) for j, v in enumerate(nu) if v > 0)

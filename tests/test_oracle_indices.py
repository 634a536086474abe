"""The oracle's multi-index machinery against brute force and the reference's known answers."""

import itertools as it
import json
import os

import numpy as np
import pytest

from oracle import indices as OI


def k_log2(d):
    return np.log([2 + i for i in range(d)]) / np.log(2)


GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                      "reference_known_answers.json")))


@pytest.mark.parametrize("case", GOLDEN["indexset_cardinality"])
def test_known_cardinalities(case):
    """tests/test_indices_runtime.py:10-12 of the reference (k_j = log2(2+j))."""
    k = k_log2(case["d"])
    assert OI.indexset_cardinality(k, case["t"]) == case["m"]
    assert len(OI.indexset(k, case["t"])) == case["m"]


@pytest.mark.parametrize("case", [c for c in GOLDEN["find_approximate_threshold"] if c["d"] <= 100])
def test_known_threshold_accuracy(case):
    """tests/test_indices_runtime.py:21-34 of the reference: the threshold's node count is within
    the requested accuracy of the target (the d = 10000 cases run on the C++ port)."""
    k = k_log2(case["d"])
    t = OI.find_approximate_threshold(k, case["m"], nested=case["nested"],
                                      accuracy=case["accuracy"])
    n = OI.nodeset_cardinality(k, t, nested=case["nested"])
    assert abs(n - case["m"]) / case["m"] < case["accuracy"]


def random_k(rng):
    d = int(rng.integers(1, 6))
    a = 1.1 + 2.9 * rng.random()
    b = 0.1 + 1.9 * rng.random()
    return np.log([a + b * i for i in range(d)]) / np.log(a)


@pytest.mark.parametrize("seed", range(8))
def test_indexset_vs_bruteforce(seed):
    """Membership over the dense box and sparse == dense sets (tests/test_indices.py:29-65)."""
    rng = np.random.default_rng(seed)
    k = random_k(rng)
    t = OI.find_approximate_threshold(k, int(rng.integers(1, 100)), nested=bool(seed % 2))
    sparse = set(OI.indexset(k, t))
    dense = OI.dense_indexset(tuple(k), t)
    assert {OI.dense_to_sparse(v) for v in dense} == sparse
    assert OI.indexset_cardinality(k, t) == len(sparse)
    for v in it.product(*[range(int(np.floor(ki)) + 2) for ki in k]):
        assert (OI.dense_to_sparse(v) in sparse) == (np.dot(v, k) < t)


@pytest.mark.parametrize("seed", range(8))
def test_zetas_vs_inclusion_exclusion(seed):
    """zeta_nu = sum_{e in {0,1}^d, nu+e in Lambda} (-1)^|e| (tests/test_indices.py:68-83)."""
    rng = np.random.default_rng(100 + seed)
    k = random_k(rng)
    t = OI.find_approximate_threshold(k, int(rng.integers(1, 100)), nested=True)
    d = len(k)
    n2nus, n2z = OI.non_zero_indices_and_zetas(k, t)
    got = {nu: z for n in n2nus for nu, z in zip(n2nus[n], n2z[n])}
    for nu in OI.indexset(k, t):
        dense = OI.sparse_to_dense(nu, d)
        z_dense = OI.dense_smolyak_coefficient(k, t, dense)
        z_sparse = OI.smolyak_coefficient(k, d, t - np.dot(dense, k), 0)
        assert z_dense == z_sparse
        assert got.get(nu, 0) == z_dense
    assert sum(got.values()) == 1  # constants are reproduced


def test_nodeset_non_nested_is_sum_of_block_sizes():
    rng = np.random.default_rng(5)
    for _ in range(5):
        k = random_k(rng)
        t = 1 + 6 * rng.random()
        n2nus, _ = OI.non_zero_indices_and_zetas(k, t)
        tot = sum(int(np.prod([v + 1 for _, v in nu])) for n in n2nus for nu in n2nus[n])
        assert OI.nodeset_cardinality(k, t, nested=False) == tot


@pytest.mark.parametrize("nested,m,acc", [(True, 1000, 0.01), (False, 1000, 0.1)])
def test_threshold_accuracy(nested, m, acc):
    """tests/test_indices_runtime.py:33-46 (d = 100)."""
    k = k_log2(100)
    t = OI.find_approximate_threshold(k, m, nested=nested, accuracy=acc)
    assert abs(OI.nodeset_cardinality(k, t, nested=nested) - m) / m < acc

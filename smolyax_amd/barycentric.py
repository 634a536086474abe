"""Barycentric weights (src/smolyax/barycentric.py:13-31) and the first-form weights the kernels use.

The reference's weights are w_b = prod_{a != b} 1/(xi_a - xi_b), which is (-1)^n times the
textbook weights; the sign cancels in the normalised basis b / sum(b).  The device kernels use the
division-free first form Lhat_mu(x) = v_mu prod_{nu != mu} (x - xi_nu) with the textbook weights
v_mu = 1 / prod_{nu != mu} (xi_mu - xi_nu); Lhat is the same Lagrange basis.
"""

import numpy as np


def compute_weights(nodes) -> np.ndarray:
    """Reference convention, barycentric.py:13-31."""
    nodes = np.asarray(nodes, dtype=np.float64)
    diffs = nodes[:, None] - nodes
    diffs = np.where(diffs == 0, 1, diffs)
    return np.prod(1 / diffs, axis=0)


def first_form_weights(nodes) -> np.ndarray:
    """v_mu = 1 / prod_{nu != mu} (xi_mu - xi_nu) for the modified Lagrange (first) form."""
    nodes = np.asarray(nodes, dtype=np.float64)
    n = len(nodes)
    v = np.empty(n)
    for m in range(n):
        p = 1.0
        for q in range(n):
            if q != m:
                p *= nodes[m] - nodes[q]
        v[m] = 1.0 / p
    return v

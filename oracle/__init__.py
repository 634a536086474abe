"""FP64 CPU oracle: a plain numpy/Python restatement of the reference smolyax algorithms.

This package is test infrastructure. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import it, and only to check or time results against.
The product package ``smolyax_amd`` never imports it and has no CPU fallback.

Modules
-------
indices   multi-index set, Smolyak coefficients, cardinalities, threshold search
          (follows src/smolyax/indices.py)
nodes     Leja / Gauss-Hermite node values, barycentric and quadrature weights
          (follows src/smolyax/nodes/leja.py, gausshermite.py, barycentric.py:13-31)
smolyak   grouped padded packing, evaluation, gradient, integral
          (follows src/smolyax/interpolation.py:140-390, barycentric.py:34-229)
"""

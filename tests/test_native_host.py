"""The native library: it loads, exports every declared symbol, and its host-side index machinery
is bit-identical to the oracle (no GPU needed)."""

import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "smolyax_amd.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(sbi_[a-z0-9_]+)\s*\(", hdr)))


def test_library_exports_all_declared_symbols():
    from smolyax_amd import _native as N

    out = subprocess.run(["nm", "-D", "--defined-only", N.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (sbi_\w+)", out))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    assert set(declared_symbols()) <= set(N.SIGNATURES), "ctypes signatures incomplete"


def test_error_reporting():
    import ctypes

    from smolyax_amd import _native as N

    plan = ctypes.c_void_p()
    rc = N.fn["sbi_bary_plan_create"](None, ctypes.byref(plan))
    assert rc == -1
    assert b"null" in N.fn["sbi_last_error"]()


NULL_CALLS = r"""
import ctypes, math, sys
sys.path.insert(0, sys.argv[1])
from smolyax_amd import _native as N
skip = {"sbi_last_error", "sbi_version", "sbi_plan_destroy", "sbi_newton_plan_destroy"}
for name, (res, args) in sorted(N.SIGNATURES.items()):
    if name in skip:
        continue
    vals = [1 if a in (ctypes.c_int64, ctypes.c_int32) else 1.0 if a is ctypes.c_double else None
            for a in args]
    rc = N.fn[name](*vals)
    print(name, "nan" if isinstance(rc, float) and math.isnan(rc) else rc,
          N.fn["sbi_last_error"]().decode().split(":")[0], flush=True)
"""


def test_every_entry_point_rejects_null_arguments():
    """Every entry point, called with NULL pointers (and counts of 1), returns its error value
    (a negative code, -1, or NaN for the threshold search) and sets sbi_last_error, without
    touching a device.  In a child process, so a crash is a failure, not a lost test run."""
    out = subprocess.run([sys.executable, "-c", NULL_CALLS, ROOT], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    rows = [line.split() for line in out.stdout.splitlines()]
    names = {r[0] for r in rows}
    assert names == set(declared_symbols()) - {"sbi_last_error", "sbi_version", "sbi_plan_destroy",
                                               "sbi_newton_plan_destroy"}
    for name, rc, who in rows:
        assert rc == "nan" or int(rc) < 0, (name, rc)
        assert who.startswith("sbi_") and who.split("_")[1] == name.split("_")[1], (name, who)


def k_rule(d):
    return np.array([np.log((2 + j) / np.log(2)) for j in range(d)])


@pytest.mark.parametrize("d,m,nested", [(10, 1000, True), (200, 5000, True), (100, 10000, True),
                                         (20, 800, False), (8, 2000, False)])
def test_native_indices_bit_exact(d, m, nested):
    from oracle import indices as OI
    from smolyax_amd import indices as SI

    k = k_rule(d)
    t_o = OI.find_approximate_threshold(k, m, nested)
    t_n = SI.find_approximate_threshold(k, m, nested)
    assert t_o == t_n  # bitwise equal double
    assert SI.nodeset_cardinality(k, t_o, nested) == OI.nodeset_cardinality(k, t_o, nested)
    assert SI.indexset_cardinality(k, t_o) == OI.indexset_cardinality(k, t_o)
    assert SI.indexset(k, t_o) == OI.indexset(k, t_o)
    a, b = SI.non_zero_indices_and_zetas(k, t_o), OI.non_zero_indices_and_zetas(k, t_o)
    assert dict(a[0]) == dict(b[0]) and dict(a[1]) == dict(b[1])
    assert list(a[0].keys()) == list(b[0].keys())


GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                      "reference_known_answers.json")))


def test_native_known_cardinalities():
    """tests/test_indices_runtime.py:10-12 of the reference, on the C++ port."""
    from smolyax_amd import indices as SI

    for case in GOLDEN["indexset_cardinality"]:
        k = np.log([2 + i for i in range(case["d"])]) / np.log(2)
        assert SI.indexset_cardinality(k, case["t"]) == case["m"]
        assert len(SI.indexset(k, case["t"])) == case["m"]


@pytest.mark.parametrize("case", GOLDEN["find_approximate_threshold"])
def test_native_known_threshold_accuracy(case):
    """tests/test_indices_runtime.py:21-34 of the reference, on the C++ port (d up to 10000)."""
    from smolyax_amd import indices as SI

    k = np.log([2 + i for i in range(case["d"])]) / np.log(2)
    t = SI.find_approximate_threshold(k, case["m"], case["nested"], accuracy=case["accuracy"])
    n = SI.nodeset_cardinality(k, t, case["nested"])
    assert abs(n - case["m"]) / case["m"] < case["accuracy"]


def test_native_smolyak_coefficient_random():
    from oracle import indices as OI
    from smolyax_amd import indices as SI

    rng = np.random.default_rng(0)
    for _ in range(50):
        d = int(rng.integers(1, 12))
        k = np.sort(rng.uniform(0.5, 3, d))
        rem = rng.uniform(0, 6)
        p = int(rng.integers(0, 2))
        assert SI.smolyak_coefficient(k, d, rem, p) == OI.smolyak_coefficient(k, d, rem, p)

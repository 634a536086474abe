"""Host side of the C ABI under AddressSanitizer and UndefinedBehaviorSanitizer (CPU, no GPU).

smolyax_amd.build.build_host_check() builds csrc/host_check.cpp against indices.cpp and the plan
entry points of bary.hip / newton.hip with ASan + UBSan on the host code.  The test program runs the
index-set entry points with exactly sized buffers and through their error paths and creates,
inspects and destroys plans (valid, incomplete and over-capacity descriptors).  Any sanitizer
report fails the run; its printed index-set results must equal the production library's.
"""

import os
import subprocess

import numpy as np

from smolyax_amd import build, indices


def test_host_abi_under_asan_and_ubsan():
    exe = build.build_host_check(verbose=False)
    env = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
    lines = r.stdout.split("\n")
    assert "ok" in lines
    cases = [ln.split() for ln in lines if ln.startswith("case ")]
    assert len(cases) == 4
    for c in cases:
        d, m, t = int(c[1]), int(c[2]), float(c[3])
        k = np.log((2 + np.arange(d)) / np.log(2))
        assert t == indices.find_approximate_threshold(k, m, True)
        assert int(c[4]) == indices.indexset_cardinality(k, t)
        lam = indices.indexset(k, t)
        flat = [(j, g) for nu in lam for j, g in nu]
        assert int(c[5]) == len(flat)
        assert int(c[6]) == sum((j + 1) * (g + 3) * (i % 7 + 1) for i, (j, g) in enumerate(flat))
        off, dims, degs, zetas = indices.nonzero_blocks(k, t)
        assert int(c[7]) == len(zetas) and int(c[8]) == int(np.sum(zetas))
        assert int(c[9]) == indices.nodeset_cardinality(k, t, nested=True)

"""Build libsmolyax_amd.so in-tree with hipcc for gfx950.

    python -m smolyax_amd.build      (or __graft_entry__.build())

The host index code is compiled with -ffp-contract=off (bit-exact budget arithmetic); the HIP
sources with packed-FP32 instructions disabled.
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libsmolyax_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["bary.hip", "newton.hip", "basis.hip", "gemm.hip", "pack.hip", "server.hip"]
# per-file flags: the set-up kernels must round every product and difference on its own (their
# results equal the host packing bit for bit)
FILE_FLAGS = {"pack.hip": ["-ffp-contract=off"]}
HOST_SOURCES = ["indices.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(objdir, src + ".o")
        _run([HIPCC, "-x", "c++", "-c", *COMMON, "-ffp-contract=off", os.path.join(CSRC, src),
              "-o", obj])
        objs.append(obj)
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, src + ".o")
        _run([HIPCC, "-c", *COMMON, *HIP_FLAGS, *FILE_FLAGS.get(src, []), path, "-o", obj])
        objs.append(obj)
    tmp = OUT + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, OUT)
    return OUT


# Host-side check of the C ABI under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).
# The sanitizer flags apply to the host compilation only (each follows -Xarch_host); the device
# code of the .hip sources is built as usual and never runs in the check.
SANITIZE = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined"]
CHECK_SOURCES = [("indices.cpp", True), ("host_check.cpp", True), ("bary.hip", False),
                 ("newton.hip", False), ("basis.hip", False), ("gemm.hip", False)]


def build_host_check(verbose=True):
    """Build smolyax_amd/_build/asan/host_check (indices.cpp, the plan entry points of bary.hip,
    newton.hip, basis.hip and gemm.hip, and the test program csrc/host_check.cpp) with ASan + UBSan on
    the host side."""
    objdir = os.path.join(HERE, "_build", "asan")
    os.makedirs(objdir, exist_ok=True)
    base = ["-O1", "-g", "-fno-omit-frame-pointer", "-std=c++17", "-Wall", "-Wno-unused-function"]
    objs = []
    for src, host_only in CHECK_SOURCES:
        obj = os.path.join(objdir, src + ".o")
        if host_only:
            cmd = [HIPCC, "-x", "c++", "-c", *base, "-ffp-contract=off", *SANITIZE]
        else:
            cmd = [HIPCC, "-c", *base, *HIP_FLAGS, *SANITIZE]
        _run(cmd + [os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    exe = os.path.join(objdir, "host_check")
    _run([HIPCC, f"--offload-arch={ARCH}", *SANITIZE, *objs, "-o", exe])
    return exe


if __name__ == "__main__":
    build()
    sys.exit(0)

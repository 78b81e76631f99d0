"""Build libsympygr_amd.so (HIP, gfx950) in-tree with hipcc.

    python -m sympygr_amd.build [--resource-usage]

The library is the product: every kernel and the C ABI of
include/sympygr_amd.h, plus the C++ drop-in bssnrhs(...) of
include/bssn_dropin.h.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libsympygr_amd.so")
ARCH = os.environ.get("SGR_OFFLOAD_ARCH", "gfx950")

SOURCES = ["bssn_kernels.hip", "bssn_capi.hip", "ghost_table.cpp", "bssnrhs_dropin.cpp"]

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 (v_pk_*_f32) is kept off for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wno-unused-command-line-argument",
]


def hipcc():
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(p):
        raise RuntimeError("hipcc not found")
    return p


def build(resource_usage=False, verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp] + srcs
    if resource_usage:
        cmd.insert(1, "-Rpass-analysis=kernel-resource-usage")
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(tmp, LIB)
    return r.stderr if resource_usage else ""


if __name__ == "__main__":
    out = build(resource_usage="--resource-usage" in sys.argv, verbose=True)
    if out:
        print(out)
    print("built", LIB)

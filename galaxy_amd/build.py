"""Build libgalaxy_bh.so (gfx950) in-tree with hipcc.

Flags: -ffp-contract=off keeps every f32/f64 operation a single IEEE
operation (no FMA contraction); correctly rounded f32 divide/sqrt; packed-FP32
code generation disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libgalaxy_bh.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GBH_OFFLOAD_ARCH", "gfx950")

SOURCES = ["bh_runtime.hip", "bh_ics.cpp"]

FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


# Host-side sanitizers for CPU test runs (no GPU): AddressSanitizer and
# UndefinedBehaviorSanitizer on the host pass only; the device code is built as
# usual.  The result goes to galaxy_amd/_san/ and is loaded when GALAXY_BH_LIB
# points there (with the clang ASan runtime preloaded).
SANITIZE_FLAGS = [
    "-Xarch_host", "-fsanitize=address",
    "-Xarch_host", "-fsanitize=undefined",
    "-Xarch_host", "-shared-libsan",
    "-Xarch_host", "-fno-omit-frame-pointer",
    "-g",
]
SAN_OUT = os.path.join(HERE, "_san", "libgalaxy_bh.so")


def build(verbose: bool = False, sanitize: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    out = SAN_OUT if sanitize else OUT
    os.makedirs(os.path.dirname(out), exist_ok=True)
    tmp = out + ".tmp"
    cmd = [HIPCC, *FLAGS, *(SANITIZE_FLAGS if sanitize else []), "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libgalaxy_bh.so")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True, sanitize="--sanitize" in sys.argv[1:]))

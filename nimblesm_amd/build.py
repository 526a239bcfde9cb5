"""Build libnimblesm_amd.so for gfx950 in-tree with hipcc.

FP64 translation units (element kernels, node kernels, host driver) are
compiled with -ffp-contract=off so that every FP64 expression keeps
NimbleSM's rounding; the mixed-precision kernels are compiled with
contraction on except where a `#pragma clang fp contract(off)` (the FP64
node update) says otherwise.  Packed FP32 instructions are disabled for all device code.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnimblesm_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("NSM_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-warning-option",
]

SOURCES = {
    "nsm_fp64.hip": ["-ffp-contract=off"],
    "nsm_node.hip": ["-ffp-contract=off"],
    # fast-honor-pragmas: FP32 element arithmetic contracts, while the FP64
    # node update keeps separate multiply and add through its
    # `#pragma clang fp contract(off)` (plain `fast` ignores that pragma)
    "nsm_mixed.hip": ["-ffp-contract=fast-honor-pragmas"],
    # the tiled element kernel's launch bound holds it at two waves per SIMD
    # (256 registers); under that bound the default scheduler beats the
    # max-ilp strategy by about 3.5%
    "nsm_tiled.hip": ["-ffp-contract=fast-honor-pragmas"],
    "nsm_tiles.cpp": ["-ffp-contract=off", "-x", "hip"],
    "nsm_model.cpp": ["-ffp-contract=off", "-x", "hip"],
    "nsm_ensemble.cpp": ["-ffp-contract=off", "-x", "hip"],
    "nsm_autotune.cpp": [],
    # surrogate network: BF16 matrix-core GEMMs, FP32 epilogues and Adam
    "nsm_mlp.hip": ["-ffp-contract=fast-honor-pragmas"],
}


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed for {cmd[-1]}")
    # keep real warnings visible, drop the host-pass note about the device-only feature flag
    for line in (r.stderr or "").splitlines():
        if "packed-fp32-ops" in line or "not a recognized feature" in line:
            continue
        if line.strip():
            sys.stderr.write(line + "\n")


def build(verbose: bool = False, out: str = LIB, extra_flags=(), obj_dir: str = OBJ) -> str:
    """Compile every translation unit and link `out` (the in-tree library by
    default; `extra_flags`, e.g. -DNSM_MIXED_MIN_WAVES=3, for tuning builds)."""
    os.makedirs(obj_dir, exist_ok=True)
    objs = []
    for src, flags in SOURCES.items():
        obj = os.path.join(obj_dir, src + ".o")
        cmd = [HIPCC, *COMMON, *flags, *extra_flags, "-c", "-o", obj, os.path.join(CSRC, src)]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

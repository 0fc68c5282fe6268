"""Builds annfmp_amd/libannf_hip.so (gfx950) from annfmp_amd/csrc with hipcc.

    python -m annfmp_amd.build

Flags:
  -ffp-contract=off   every FP32 distance keeps one rounding per operation,
                      the arithmetic of the reference's openmp build;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions are generated.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libannf_hip.so")
SOURCES = ["annf_kernels.hip", "annf_eval.hip", "annf_leafmm.hip", "annf_mlp.hip", "annf_sort.hip", "annf_abi.hip"]
ARCH = os.environ.get("ANNF_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False, out=OUT):
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I", os.path.join(ROOT, "include"), "-I", SRC,
           "-o", out + ".tmp"] + [os.path.join(SRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libannf_hip.so")
    if verbose and res.stderr.strip():
        # the host pass warns about the device-only target feature; keep other warnings visible
        lines = [ln for ln in res.stderr.splitlines() if "packed-fp32-ops" not in ln]
        if lines:
            sys.stderr.write("\n".join(lines) + "\n")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))

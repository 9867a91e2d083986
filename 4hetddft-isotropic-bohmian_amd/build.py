"""Build libhetddft.so in place with hipcc for gfx950.

    python 4hetddft-isotropic-bohmian_amd/build.py
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "csrc" / "hetddft.hip"
OUT = HERE / "libhetddft.so"

HIPCC_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
    # keep the reference's separate multiply/add rounding in FP64 paths
    "-ffp-contract=off",
    # no packed FP32 instructions (build notes)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    deps = sorted((HERE / "csrc").glob("*")) + [HERE.parent / "include" / "hetddft.h"]
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", str(tmp), str(SRC)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode}) on {SRC}")
    os.replace(tmp, OUT)
    _ = deps
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))

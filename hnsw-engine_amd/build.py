"""Builds libhnsw_mi355x.so (HIP kernels for gfx950 + the C ABI + the host graph builder).

Run ``python hnsw-engine_amd/build.py``; the library lands next to this file's package
(``hnsw_amd/libhnsw_mi355x.so``), in-tree, so it travels with the repository snapshot.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "hnsw_amd" / "libhnsw_mi355x.so"
ARCH = os.environ.get("HM_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> Path:
    sources = sorted(str(p) for p in CSRC.glob("*.hip")) + sorted(str(p) for p in CSRC.glob("*.cpp"))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [
        hipcc(),
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-shared",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-result",
        "-o", str(tmp),
        *sources,
    ]
    if verbose:
        print(" ".join(cmd))
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("building libhnsw_mi355x.so failed")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))

"""Builds libhnsw_mi355x.so (HIP kernels for gfx950 + the C ABI + the host graph builder).

Run ``python hnsw-engine_amd/build.py``; the library lands next to this file's package
(``hnsw_amd/libhnsw_mi355x.so``), in-tree, so it travels with the repository snapshot.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor
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


FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _compile(src: str, obj: str, verbose: bool) -> None:
    cmd = [hipcc(), *FLAGS, "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd))
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"compiling {Path(src).name} failed")


def build(verbose: bool = False) -> Path:
    """Compiles every source to an object in parallel (one hipcc per file), then links."""
    sources = sorted(str(p) for p in CSRC.glob("*.hip")) + sorted(str(p) for p in CSRC.glob("*.cpp"))
    tmp = OUT.with_suffix(".so.tmp")
    workers = max(1, min(len(sources), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)), 16))
    with tempfile.TemporaryDirectory(prefix="hm_build_") as objdir:
        objs = [str(Path(objdir) / (Path(src).name + ".o")) for src in sources]
        # The slowest files first, so the pool ends together.
        order = sorted(range(len(sources)), key=lambda i: -Path(sources[i]).stat().st_size)
        with ThreadPoolExecutor(max_workers=workers) as pool:
            futures = [pool.submit(_compile, sources[i], objs[i], verbose) for i in order]
            for f in futures:
                f.result()
        cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs]
        if verbose:
            print(" ".join(cmd))
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError("linking libhnsw_mi355x.so failed")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))

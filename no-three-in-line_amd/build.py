"""Builds libno3.so (the gfx950 HIP engine) in-tree with hipcc.

Each translation unit is compiled to an object in parallel, then linked into
a shared library next to this file.  The device code is built for gfx950
only; packed FP32 instructions are disabled (-packed-fp32-ops) as required on
the target machines.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libno3.so"
OBJ = HERE / "_build"
SOURCES = ["runtime.hip", "score.hip", "delta.hip", "temper.hip", "mlp.hip", "learned.hip", "search.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _compile(src: str) -> Path:
    obj = OBJ / (Path(src).stem + ".o")
    cmd = [HIPCC, *COMMON, "-c", str(CSRC / src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stderr}")
    # the host pass does not know the packed-fp32 feature and says so; keep
    # everything else visible
    noise = "'-packed-fp32-ops' is not a recognized feature"
    rest = "\n".join(line for line in res.stderr.splitlines() if noise not in line and line.strip())
    if rest.strip():
        sys.stderr.write(rest + "\n")
    return obj


def build(verbose: bool = False) -> Path:
    """Compile every HIP source and link libno3.so (always rebuilds)."""
    if not Path(HIPCC).exists() and shutil.which(HIPCC) is None:
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    OBJ.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 1)) as pool:
        objs = list(pool.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,-z,defs", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)

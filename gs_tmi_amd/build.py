"""Build recipe for the gfx950 shared libraries of this package.

``python -m gs_tmi_amd.build`` compiles every HIP library in-tree (next to
this file, under ``_lib/``) with hipcc.  The libraries travel with the source
tree; nothing is installed into site-packages.
"""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
REPO_DIR = PKG_DIR.parent
LIB_DIR = PKG_DIR / "_lib"
CSRC_DIR = PKG_DIR / "csrc"
INCLUDE_DIR = REPO_DIR / "include"

ARCH = "gfx950"

# -packed-fp32-ops: keep hipcc from forming v_pk_{add,mul,fma}_f32.
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]

# (library name, sources) pairs.
LIBRARIES = [
    ("libgstv.so", ["gstv.hip"]),
    ("libgstv_mlp.so", ["mlp.hip"]),
]


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = Path(rocm) / "bin" / "hipcc"
    return str(cand) if cand.exists() else "hipcc"


def build_library(name: str, sources: list[str], verbose: bool = False) -> Path:
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    out = LIB_DIR / name
    tmp = LIB_DIR / (name + ".tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, f"-I{INCLUDE_DIR}", *[str(CSRC_DIR / s) for s in sources], "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature; that warning is harmless.
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in res.stderr.splitlines() if noise not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {name} (exit {res.returncode}):\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, out)
    return out


def build_all(verbose: bool = False) -> list[Path]:
    return [build_library(name, srcs, verbose) for name, srcs in LIBRARIES]


if __name__ == "__main__":
    for p in build_all(verbose=True):
        print(p)

"""Compile the gfx950 HIP library ``libexactdiag_hip.so`` in-tree with hipcc."""

from __future__ import annotations

import os
import pathlib
import subprocess

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libexactdiag_hip.so"
SOURCES = ["ed_sector.hip", "ed_matrix.hip", "ed_hv.hip", "ed_hmat.hip", "ed_batch.hip", "ed_lanczos.hip", "ed_mlp.hip", "ed_csr.hip", "ed_multik.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-shared",
    # keep every product separately rounded: the matrix-element values are
    # reproduced bit for bit
    "-ffp-contract=off",
    # no packed FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return str(pathlib.Path(rocm) / "bin" / "hipcc")


def build_library(verbose: bool = False) -> pathlib.Path:
    """Compile all HIP sources into one shared library (always recompiles)."""
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, f"-I{ROOT / 'include'}", f"-I{CSRC}", *[str(CSRC / s) for s in SOURCES], "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose=True))

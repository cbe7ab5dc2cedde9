"""Compile the gfx950 library in-tree (``snncuda_amd/libsnncuda_hip.so``) with hipcc.

Flags: FP contraction off and denormals kept, so FP32 results round exactly as the reference's
baseline-x86-64 build does; packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import pathlib
import shutil
import subprocess

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
SOURCES = [PKG / "csrc" / "snn_step.hip", PKG / "csrc" / "snn_adjoint.hip", PKG / "csrc" / "snn_readout.hip"]
HEADERS = sorted((PKG / "csrc").glob("*.h")) + [ROOT / "include" / "snncuda_hip.h"]
OUT = PKG / "libsnncuda_hip.so"

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and pathlib.Path(cand).exists():
            return cand
    raise FileNotFoundError("hipcc not found")


def build(verbose: bool = False) -> pathlib.Path:
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", str(tmp)] + [str(s) for s in SOURCES]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stderr[-8000:])
    if verbose and res.stderr:
        print(res.stderr[-2000:])
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))

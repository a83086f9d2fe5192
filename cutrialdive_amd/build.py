"""Builds the native parts in-tree.

  cutrialdive_amd/libctd_hip.so   HIP engine (gfx950) behind include/ctd_hip.h
  cutrialdive_amd/cutrialdive     C++ CLI linked against it (the drop-in binary)

Run as `python -m cutrialdive_amd.build` or through __graft_entry__.build().
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed FP32 instructions stay disabled in every HIP compilation
HIP_FLAGS = [
    "-O3",
    f"--offload-arch={ARCH}",
    "-std=c++20",
    "-fPIC",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    # the FP64 residue forms rely on exact Dekker products: no implicit fma contraction
    "-ffp-contract=off",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]

LIB = os.path.join(HERE, "libctd_hip.so")
CLI = os.path.join(HERE, "cutrialdive")


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib() -> str:
    srcs = [os.path.join(CSRC, "ctd_engine.hip")]
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *srcs])
    os.replace(tmp, LIB)
    return LIB


def build_cli() -> str | None:
    if not os.path.isdir(HOST):
        return None
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp"))
    if not srcs:
        return None
    tmp = CLI + ".tmp"
    _run(
        [
            "g++",
            "-O2",
            "-std=c++20",
            f"-I{INCLUDE}",
            f"-I{HOST}",
            "-o",
            tmp,
            *srcs,
            f"-L{HERE}",
            "-lctd_hip",
            "-Wl,-rpath,$ORIGIN",
            "-lpthread",
        ]
    )
    os.replace(tmp, CLI)
    return CLI


def build_all() -> None:
    build_lib()
    build_cli()


if __name__ == "__main__":
    build_all()
    sys.exit(0)

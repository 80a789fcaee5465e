"""Build the native pieces in-tree.

    libsdr.so          HIP decode library (gfx950) behind include/sdr/decode.h
    libsdr_encode.so   DuckDB-format encoders and seeded generators (host)
    oracle/libsdr_oracle.so   CPU oracle (test infrastructure, built from oracle/)

Run as ``python -m sirius_amd.build``.  Every call recompiles.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
ARCH = os.environ.get("SDR_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",            # ALP products must not be fused
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]

HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra"]


def _run(cmd: list[str]) -> None:
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def hip_sources() -> list[Path]:
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def build_hip() -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = hip_sources()
    objs = [BUILD / (s.name + ".o") for s in srcs]

    def compile_one(pair):
        src, obj = pair
        _run([HIPCC, *HIP_FLAGS, "-x", "hip", "-c", str(src), "-o", str(obj)])

    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        list(ex.map(compile_one, zip(srcs, objs)))
    out = PKG / "libsdr.so"
    tmp = PKG / "libsdr.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, out)
    return out


def build_encoder() -> Path:
    out = PKG / "libsdr_encode.so"
    tmp = PKG / "libsdr_encode.so.tmp"
    _run(["g++", *HOST_FLAGS, "-shared", "-o", str(tmp), str(CSRC / "encode" / "sdr_encode.cpp")])
    os.replace(tmp, out)
    return out


def build_oracle() -> Path:
    _run(["make", "-s", "-B", "-C", str(ROOT / "oracle")])
    return ROOT / "oracle" / "libsdr_oracle.so"


def build_all() -> None:
    build_oracle()
    build_encoder()
    build_hip()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "encoder": build_encoder, "oracle": build_oracle}[w]()

"""Builds libibfs.so (HIP for gfx950 + host C++) in-tree with hipcc.

Usage: python -m ibfs_amd.build   (or ibfs_amd.build.build_native()).
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "libibfs.so"
CLI = PKG / "gpu-ibfs"
CONVERTER = PKG / "text-to-bin"

HIP_SOURCES = ["kernels_traverse.hip", "kernels_emit.hip", "engine.hip", "batch.hip",
               "serve.hip", "noalloc.hip", "host_copy.hip"]
CXX_SOURCES = ["hostcsr.cpp", "autotune.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 forms are disabled for all device code on this target.
DEVICE_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
                "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def build_native(verbose: bool = False) -> Path:
    objdir = ROOT / "_build" / "obj"
    objdir.mkdir(parents=True, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        if not (CSRC / src).exists():
            continue
        obj = objdir / (src + ".o")
        _run([HIPCC, *COMMON, *DEVICE_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(str(obj))
    for src in CXX_SOURCES:
        obj = objdir / (src + ".o")
        _run([HIPCC, *COMMON, "-x", "c++", "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(str(obj))
    tmp = OUT.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", *DEVICE_FLAGS, "-o", str(tmp), *objs, "-lpthread"])
    os.replace(tmp, OUT)
    # reference-style command line front end over the C ABI
    cli_src = CSRC / "cli_main.cpp"
    if cli_src.exists():
        _run([HIPCC, *COMMON, "-x", "c++", str(cli_src), "-o", str(CLI), f"-L{PKG}", "-libfs",
              f"-Wl,-rpath,$ORIGIN", "-lpthread"])
    # the converter's command line over the C ABI
    conv_src = CSRC / "text_to_bin_main.cpp"
    if conv_src.exists():
        _run([HIPCC, *COMMON, "-x", "c++", str(conv_src), "-o", str(CONVERTER), f"-L{PKG}",
              "-libfs", f"-Wl,-rpath,$ORIGIN", "-lpthread"])
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build_native(verbose=True)

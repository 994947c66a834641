"""Ahead-of-time build of the gfx950 library and CLI, in-tree.

    python -m pfcsieve_amd.build

produces ``pfcsieve_amd/libpfcsieve.so`` (HIP kernels + C ABI of
``include/pfc.h``) and ``pfcsieve_amd/pfcsieve`` (the command-line program).
Both are compiled with hipcc for gfx950 only; nothing is compiled at run time.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libpfcsieve.so")
CLI = os.path.join(PKG, "pfcsieve")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = "gfx950"
# Packed-FP32 instruction selection is disabled for all device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC", "-Wall",
          "-Wno-unused-command-line-argument", *DEVICE_FLAGS]

LIB_SOURCES = ["pfc_kernels.hip", "pfc_host.cpp"]
CLI_SOURCES = ["pfcsieve_main.cpp"]


def _run(cmd):
    proc = subprocess.run(cmd, cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"build failed: {' '.join(cmd)}")
    return out


def build_library() -> str:
    srcs = [os.path.join(CSRC, s) for s in LIB_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, *COMMON, "-shared", *srcs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_cli() -> str:
    srcs = [os.path.join(CSRC, s) for s in CLI_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if not srcs:
        return ""
    tmp = CLI + ".tmp"
    _run([HIPCC, *COMMON, *srcs, "-L" + PKG, "-lpfcsieve", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, CLI)
    return CLI


def build_all() -> None:
    build_library()
    build_cli()


if __name__ == "__main__":
    build_all()
    print(LIB)

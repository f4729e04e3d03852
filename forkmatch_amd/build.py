"""Builds libforkmatch.so and the ForkMatch CLI in-tree for gfx950, and the
CPU oracle library (test infrastructure) under oracle/.

Run as ``python -m forkmatch_amd.build`` or through ``__graft_entry__.build()``.
hipcc cross-compiles the device code without a GPU.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
LIB = PKG / "libforkmatch.so"
CLI = PKG / "ForkMatch"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liboracle.so"

ARCH = os.environ.get("FM_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 instructions are disabled for every device compile (the engine
# has no floating point, this only guards against generated code using them).
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True)
    noise = "'-packed-fp32-ops' is not a recognized feature"
    err = "\n".join(l for l in proc.stderr.splitlines() if noise not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{err}\n{proc.stdout}")
    if err.strip():
        sys.stderr.write(err + "\n")


def build_library() -> Path:
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    common = ["-O3", "-std=c++17", "-fPIC", f"-I{INCLUDE}", f"-I{CSRC}", "-Wall", "-Wno-unused-function"]
    host_obj = objdir / "host.o"
    engine_obj = objdir / "engine.o"
    gbuild_obj = objdir / "gbuild.o"
    isect_obj = objdir / "isect.o"
    _run([HIPCC, "-x", "c++", *common, "-c", str(CSRC / "host.cpp"), "-o", str(host_obj)])
    _run([HIPCC, f"--offload-arch={ARCH}", *common, *DEVICE_FLAGS, "-c", str(CSRC / "engine.hip"),
          "-o", str(engine_obj)])
    _run([HIPCC, f"--offload-arch={ARCH}", *common, *DEVICE_FLAGS, "-c", str(CSRC / "gbuild.hip"),
          "-o", str(gbuild_obj)])
    _run([HIPCC, f"--offload-arch={ARCH}", *common, *DEVICE_FLAGS, "-c", str(CSRC / "isect.hip"),
          "-o", str(isect_obj)])
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), str(host_obj),
          str(engine_obj), str(gbuild_obj), str(isect_obj), "-lpthread"])
    os.replace(tmp, LIB)
    tmp_cli = CLI.with_suffix(".tmp")
    _run([HIPCC, "-x", "c++", *common, str(CSRC / "cli.cpp"), "-o", str(tmp_cli), f"-L{PKG}", "-lforkmatch",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp_cli, CLI)
    return LIB


def build_oracle() -> Path:
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    _run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-Wall", "-o", str(tmp), str(ORACLE_DIR / "oracle.c"),
          "-lpthread"])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def build() -> None:
    build_oracle()
    build_library()


if __name__ == "__main__":
    build()
    print(f"built {LIB}, {CLI}, {ORACLE_LIB}")

"""Builds the unicycle plugin test library (tests/plugin/_build/libunicycle_plugin.so): the plugin translation unit
with hipcc for gfx950 (the way a user compiles a plugin against include/mppi_amd/device_plugin.hpp), the CPU
reference with g++ -ffp-contract=off, linked against libmppi_amd.so."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
OUT = HERE / "_build"
LIB = OUT / "libunicycle_plugin.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
PKG = ROOT / "mppi_generic_vendor_amd"


def build(verbose: bool = False) -> Path:
    OUT.mkdir(parents=True, exist_ok=True)
    hip_obj = OUT / "unicycle_plugin.o"
    cpu_obj = OUT / "unicycle_oracle.o"
    cmds = [
        [HIPCC, "-std=c++17", "-O3", "-fPIC", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-c", str(HERE / "unicycle_plugin.hip"), "-o", str(hip_obj)],
        ["g++", "-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-c", str(HERE / "unicycle_oracle.cpp"), "-o",
         str(cpu_obj)],
    ]
    for cmd in cmds:
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"plugin test build failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    tmp = LIB.with_suffix(".so.tmp")
    link = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", str(hip_obj), str(cpu_obj), "-o", str(tmp),
            f"-L{PKG}", "-lmppi_amd", "-Wl,-rpath,$ORIGIN/../../../mppi_generic_vendor_amd", "-Wl,-rpath,/opt/rocm/lib"]
    res = subprocess.run(link, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"plugin test link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)

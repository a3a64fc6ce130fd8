"""Builds the gfx950 engine (libsdbfx.so) in-tree with hipcc.

    python -m sdhash_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsdbfx.so")
CXXLIB = os.path.join(HERE, "libsdbf_amd.so")
CLI = os.path.join(HERE, "sdhash-amd")
SYNTH_LIB = os.path.join(HERE, "libsdbfx_synth.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # IEEE double on the device and on the host: no contraction into FMA
    "-ffp-contract=off",
    # keep hipcc from emitting packed fp32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd, verbose):
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = ("packed-fp32-ops",)
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if not any(n in l for n in noise))
    if r.returncode != 0:
        raise RuntimeError(f"{cmd[0]} failed:\n" + msg)
    if verbose and msg.strip():
        print(msg, file=sys.stderr)


def build(verbose: bool = False) -> str:
    # the engine: kernels + C ABI
    _run([HIPCC, *FLAGS, "-I", INCLUDE, "-x", "hip", os.path.join(CSRC, "sdbfx_host.cpp"),
          os.path.join(CSRC, "sdbfx_serve.cpp"), "-o", LIB + ".tmp"], verbose)
    os.replace(LIB + ".tmp", LIB)
    # the device-side synthetic input generator used by bench.py
    _run([HIPCC, *FLAGS, "-x", "hip", os.path.join(CSRC, "synth_fill.hip"), "-o", SYNTH_LIB + ".tmp"], verbose)
    os.replace(SYNTH_LIB + ".tmp", SYNTH_LIB)
    # the C++ sdbf / sdbf_set classes over the C ABI, and the sdhash CLI
    cxx = ["g++", "-std=c++17", "-O2", "-fPIC", "-Wall", "-I", INCLUDE]
    _run([*cxx, "-shared", os.path.join(CSRC, "sdbf_api.cpp"), os.path.join(CSRC, "sdhash_srv.cpp"), "-o",
          CXXLIB + ".tmp", "-L", HERE,
          "-l:libsdbfx.so", "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(CXXLIB + ".tmp", CXXLIB)
    _run([*cxx, os.path.join(CSRC, "sdhash_cli.cpp"), "-o", CLI + ".tmp", "-L", HERE, "-l:libsdbf_amd.so",
          "-l:libsdbfx.so", "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(CLI + ".tmp", CLI)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))

"""Build the native libraries in-tree.

    python -m radiolensfit2_amd.build        # librl2hip.so (gfx950) + librl2sim.so

librl2hip.so: hand-written HIP kernels for gfx950 plus the host runtime, fit driver and
reference-signature drop-ins, linked into one shared library with a plain C ABI
(include/rl2.h, include/rl2_reference.h).  librl2sim.so: host-only synthetic observation
and catalog generators (include/rl2_sim.h).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
HIP_LIB = os.path.join(PKG, "librl2hip.so")
SIM_LIB = os.path.join(PKG, "librl2sim.so")

ARCH = os.environ.get("RL2_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-std=c++20", "-fPIC", "-ffp-contract=off", f"-I{INC}", f"-I{CSRC}",
          "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]
# Device code: hardware FP atomics on coarse-grained memory; no packed-FP32 instructions.
DEVICE = [f"--offload-arch={ARCH}", "-munsafe-fp-atomics",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

HIP_SOURCES = ["rl2_kernels.hip", "rl2_runtime.hip", "rl2_mlp.hip"]
HOST_SOURCES = ["rl2_fit.cpp", "rl2_dropin.cpp", "rl2_proposal.cpp"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_sim():
    cxx = shutil.which("g++") or "g++"
    _run([cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
          f"-I{INC}", os.path.join(CSRC, "rl2_sim.cpp"), "-o", SIM_LIB + ".tmp"])
    os.replace(SIM_LIB + ".tmp", SIM_LIB)


def build_hip(jobs: int = 4):
    os.makedirs(BUILD, exist_ok=True)
    objs, procs = [], []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        cmd = [HIPCC, "-x", "hip", *COMMON, *DEVICE, "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd)))
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        cmd = [HIPCC, "-x", "hip", *COMMON, *DEVICE, "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd)))
        objs.append(obj)
    failed = False
    for cmd, p in procs:
        if p.wait() != 0:
            print("FAILED:", " ".join(cmd), file=sys.stderr)
            failed = True
    if failed:
        raise RuntimeError("librl2hip build failed")
    _run([HIPCC, *DEVICE, "-shared", "-fPIC", *objs, "-o", HIP_LIB + ".tmp"])
    os.replace(HIP_LIB + ".tmp", HIP_LIB)


def build_all():
    build_sim()
    build_hip()


if __name__ == "__main__":
    build_all()

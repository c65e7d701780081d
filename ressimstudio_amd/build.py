"""Build the native libraries in-tree (they travel with the source tree; nothing is installed).

  libressim_hip.so   the product: HIP kernels for gfx950 + host set-up + C-ABI (hipcc)
  libressim_gen.so   the seeded workload generator and engine-side well Schur reduction (g++, OpenMP)
  host_check         the host set-up (level schedules, colouring, slices, AMG set-up, well Schur plan) under the host
                     AddressSanitizer / UndefinedBehaviorSanitizer (g++; tests/test_host_sanitized.py runs it)
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
HIP_SOURCES = ["csrc/hip/rsh_kernels.hip", "csrc/hip/rsh_setup.cpp", "csrc/hip/rsh_api.cpp", "csrc/hip/rsh_batch.cpp", "csrc/hip/rsh_server.cpp", "csrc/hip/rsh_wells.cpp"]
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
             # packed FP32 instructions stay disabled (hipcc's host pass warns about the feature: harmless)
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=HERE, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stdout.write(r.stdout)
        raise RuntimeError(f"build failed: {' '.join(cmd)}")


def build_hip():
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    _run([hipcc, *HIP_FLAGS, "-o", "libressim_hip.so", *HIP_SOURCES])


def build_gen():
    cxx = os.environ.get("CXX", "g++")
    _run([cxx, "-std=c++17", "-O2", "-fopenmp", "-fPIC", "-shared", "-o", "libressim_gen.so", "csrc/gen/rs_gen.cpp"])


# the host set-up alone (no device code) under the host sanitizers: g++, never hipcc
HOST_CHECK_FLAGS = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-D__HIP_PLATFORM_AMD__",
                    "-I/opt/rocm/include"]


def build_host_check():
    cxx = os.environ.get("CXX", "g++")
    _run([cxx, *HOST_CHECK_FLAGS, "-o", "host_check", "csrc/host/host_check.cpp", "csrc/hip/rsh_setup.cpp"])


def build_all():
    build_gen()
    build_hip()
    build_host_check()


if __name__ == "__main__":
    build_all()

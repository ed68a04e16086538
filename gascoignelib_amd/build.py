"""Builds the gascoignelib_amd native libraries in-tree.

  libgascoigne_hip.so   HIP kernels + C ABI of include/gascoigne_hip.h (gfx950)
  libgh_structure.so    host structure producer, include/gh_structure.h

Run as `python -m gascoignelib_amd.build` or through __graft_entry__.build().
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no packed FP32 instructions: some of their forms have given wrong results
    # on this hardware without any error being raised
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]

HIP_SOURCES = ["gh_core.hip", "gh_solver.hip", "gh_adjoint.hip", "gh_vanka.hip", "gh_assemble.hip", "gh_ilu.hip",
               "gh_dwr.hip", "gh_direct.hip"]
HIP_LIBS = ["-lrocsolver", "-lrocblas"]
HIP_LIB = os.path.join(PKG, "libgascoigne_hip.so")

CXX = os.environ.get("CXX", "g++")
STRUCT_SOURCES = ["structure.cc"]
STRUCT_LIB = os.path.join(PKG, "libgh_structure.so")


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_hip():
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    tmp = HIP_LIB + ".tmp"
    run([HIPCC] + HIP_FLAGS + srcs + HIP_LIBS + ["-o", tmp])
    os.replace(tmp, HIP_LIB)


def build_structure():
    srcs = [os.path.join(CSRC, s) for s in STRUCT_SOURCES]
    tmp = STRUCT_LIB + ".tmp"
    run([CXX, "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-Wall",
         f"-I{os.path.join(ROOT, 'include')}"] + srcs + ["-o", tmp])
    os.replace(tmp, STRUCT_LIB)


def main():
    build_structure()
    build_hip()


if __name__ == "__main__":
    sys.exit(main())

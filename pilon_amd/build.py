"""Compile the native parts of pilon_amd in-tree.

libpilonhip.so       gfx950 kernels + C ABI (hipcc --offload-arch=gfx950)
libpl_core_host.so   host build of the kernel core, used by CPU tests only
libpilon_synth.so    synthetic workload generator (host C++, OpenMP)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_engine():
    _run(["hipcc", *HIP_FLAGS, "-o", os.path.join(HERE, "libpilonhip.so"), os.path.join(CSRC, "pl_engine.hip")])


def build_core_host():
    _run(["hipcc", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", os.path.join(HERE, "libpl_core_host.so"),
          os.path.join(CSRC, "core_host.cpp")])


def build_synth():
    _run(["g++", "-O2", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-o", os.path.join(HERE, "libpilon_synth.so"),
          os.path.join(HERE, "synth", "synth.cpp")])


def build_all():
    build_engine()
    build_core_host()
    build_synth()


if __name__ == "__main__":
    build_all()
    sys.exit(0)

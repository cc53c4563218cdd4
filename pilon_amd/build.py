"""Compile the native parts of pilon_amd in-tree.

libpilonhip.so       gfx950 kernels + C ABI (hipcc --offload-arch=gfx950)
libpl_core_host.so   host build of the kernel core, used by CPU tests only
libpilon_synth.so    synthetic workload generator (host C++, OpenMP)
libpilonbam.so       BAM reader producing read batches (host C++, libdeflate, threads)
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
          os.path.join(HERE, "synth", "synth.cpp"), "-lz"])


def build_bam():
    # libdeflate: header and link library of the image's conda tree; the
    # shared object (libdeflate.so.0) is also in the system library path
    _run(["g++", "-O3", "-std=c++17", "-pthread", "-fPIC", "-shared", "-I/opt/conda/include",
          "-o", os.path.join(HERE, "libpilonbam.so"), os.path.join(CSRC, "pl_bam.cpp"),
          "/opt/conda/lib/libdeflate.so"])


def build_all():
    build_engine()
    build_core_host()
    build_synth()
    build_bam()


if __name__ == "__main__":
    build_all()
    sys.exit(0)

"""Builds the native libraries in-tree.

  lib/libculsh.so           hipcc, gfx950: the LSH index/query library (C ABI
                            in include/culsh_c.h)
  lib/libculsh_workload.so  g++: seeded synthetic workload generator

Run ``python -m culsh_amd.build`` or ``__graft_entry__.build()``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    # Packed FP32 (v_pk_*_f32) stays off: ordinary float code only.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    # __launch_bounds__ occupancy targets are hints; paths whose registers do
    # not fit them run at the occupancy they get.
    "-Wno-pass-failed",
]
CXX_FLAGS = ["-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared",
             "-std=c++17", "-Wall"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_workload():
    os.makedirs(LIBDIR, exist_ok=True)
    _run(["g++", *CXX_FLAGS, "-o", os.path.join(LIBDIR, "libculsh_workload.so"),
          os.path.join(CSRC, "workload.cpp")])


def build_culsh():
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libculsh.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp,
          os.path.join(CSRC, "culsh.hip")])
    os.replace(tmp, out)


def build_cli():
    """culsh_cli: the reference CLI's contract, a plain C++ caller of cuLSH.h."""
    os.makedirs(LIBDIR, exist_ok=True)
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
          os.path.join(CSRC, "cli.cpp"), "-L", LIBDIR, "-lculsh",
          f"-Wl,-rpath,$ORIGIN", "-o", os.path.join(LIBDIR, "culsh_cli")])


def build_all():
    build_workload()
    build_culsh()
    build_cli()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "culsh": build_culsh, "workload": build_workload, "cli": build_cli}[w]()

"""Builds libbankai_gpu.so in-tree: HIP kernels for gfx950 with hipcc, the host
learner and dataset construction with g++, linked against the HIP runtime.

All floating-point code is compiled with -ffp-contract=off: the reference's
split arithmetic is unfused on x86_64 and the quantized path must reproduce it
bit for bit. -packed-fp32-ops is disabled for the device code.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbankai_gpu.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-Wall",
             "-Wno-unused-parameter", "-fopenmp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]))
    # keep warnings visible except the known host-pass notice about packed-fp32-ops
    err = "\n".join(line for line in r.stderr.splitlines() if "packed-fp32-ops" not in line)
    if err.strip() and verbose:
        sys.stderr.write(err + "\n")


def build(verbose=False):
    """Compiles every source and links the shared library (always rebuilds)."""
    os.makedirs(BUILD, exist_ok=True)
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    objs = []
    k_obj = os.path.join(BUILD, "kernels.o")
    _run([hipcc, *HIP_FLAGS, "-c", os.path.join(CSRC, "kernels.hip"), "-o", k_obj], verbose)
    objs.append(k_obj)
    for src in ("learner.cpp", "dataset.cpp"):
        obj = os.path.join(BUILD, src.replace(".cpp", ".o"))
        _run(["g++", *CXX_FLAGS, "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"),
              "-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run(["g++", "-shared", "-o", tmp, *objs, "-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-fopenmp",
          "-Wl,-rpath," + os.path.join(ROCM, "lib")], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))

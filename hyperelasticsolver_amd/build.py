"""Build step for libhes.so (gfx950) and the CPU oracle.

`build_lib()` compiles every HIP translation unit of the package into one
shared library next to this file; it is what `__graft_entry__.build()` runs.
"""
import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libhes.so")

SOURCES = ["hes_kernels.hip", "hes_batch.hip", "hes_adjoint.hip", "hes_badjoint.hip",
           "hes_noalloc.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # never emit packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # multiply-adds are fused only within one source expression: every kernel
    # instantiation that shares a routine (pinned or not, solver or batch)
    # then rounds it the same way, whatever its control flow
    "-ffp-contract=on",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + CSRC,
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def build_lib(extra_flags=(), verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + list(extra_flags) + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


if __name__ == "__main__":
    build_lib(verbose=True)
    build_oracle()

"""Build librcgpu.so in-tree with hipcc for gfx950.

One hipcc invocation compiles the HIP kernels and the host C++ into a shared library with no PyTorch dependency.
-ffp-contract=off keeps every f32 operation of the DCT path separately rounded (the SSIMULACRA2 kernels opt back
into contraction per scope); -Xclang -target-feature -Xclang -packed-fp32-ops keeps packed FP32 instructions out of
the generated code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librcgpu.so")
SOURCES = ["api.cpp", "batch.cpp", "jpeg.hip", "ssim2.hip", "synth.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RCG_OFFLOAD_ARCH", "gfx950")


def command(out=LIB):
    return [HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-result",
            "-o", out] + [os.path.join(CSRC, s) for s in SOURCES] + ["-lpthread"]


def build(verbose=False):
    cmd = command()
    tmp = LIB + ".tmp"
    cmd[cmd.index("-o") + 1] = tmp
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)

"""Build librockfit_dem.so for gfx950 with hipcc (in-tree, lib/).

Flags: -O3, FP64 operation order preserved (-ffp-contract=off), packed FP32
disabled (-packed-fp32-ops).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = ["csrc/rf_kernels.hip", "csrc/rf_runtime.hip"]
OUT = os.path.join(HERE, "lib", "librockfit_dem.so")
ARCH = os.environ.get("RF_OFFLOAD_ARCH", "gfx950")


def hipcc_cmd():
    return ["hipcc", "-O3", f"--offload-arch={ARCH}", "-fPIC", "-shared", "-std=c++17",
            "-ffp-contract=off", "-Wno-unused-result", "-Wno-unused-value",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I" + os.path.join(ROOT, "include"), "-o", OUT] + SOURCES


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = hipcc_cmd()
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, cwd=HERE, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc build of librockfit_dem.so failed")
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))

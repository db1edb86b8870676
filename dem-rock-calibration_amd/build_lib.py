"""Build librockfit_dem.so for gfx950 with hipcc (in-tree, lib/).

Flags: -O3, FP64 operation order preserved (-ffp-contract=off), packed FP32
disabled (-packed-fp32-ops).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = ["csrc/rf_kernels.hip", "csrc/rf_runtime.hip", "csrc/rf_fire.hip", "csrc/rf_sens.hip"]
OUT = os.path.join(HERE, "lib", "librockfit_dem.so")
ARCH = os.environ.get("RF_OFFLOAD_ARCH", "gfx950")


def hipcc_cmd(out=OUT, defines=()):
    return (["hipcc", "-O3", f"--offload-arch={ARCH}", "-fPIC", "-shared", "-std=c++17",
             "-ffp-contract=off", "-Wno-unused-result", "-Wno-unused-value",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include")] + [f"-D{d}" for d in defines] +
            ["-o", out] + SOURCES)


def build(verbose=False, out=OUT, defines=()):
    """Compile the library; `defines` are tuning macros of rf_kernels.hip."""
    out = os.path.abspath(out)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    cmd = hipcc_cmd(out, defines)
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, cwd=HERE, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc build of librockfit_dem.so failed")
    return out


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=OUT)
    ap.add_argument("-D", dest="defines", action="append", default=[])
    a = ap.parse_args()
    print(build(verbose=True, out=a.out, defines=a.defines))

"""Build librockfit_dem.so for gfx950 with hipcc (in-tree, lib/) and the
RockFit-DEM command-line drop-in over it (bin/).

Flags: -O3, FP64 operation order preserved (-ffp-contract=off), packed FP32
disabled (-packed-fp32-ops).  Each source is compiled on its own, the step
kernels' file with LLVM's max-ilp instruction scheduler (0.8 % faster step
kernel, the same registers), then the objects are linked.  The command line
is host code only (g++).
"""
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = ["csrc/rf_kernels.hip", "csrc/rf_runtime.hip", "csrc/rf_fire.hip", "csrc/rf_sens.hip"]
OUT = os.path.join(HERE, "lib", "librockfit_dem.so")
ARCH = os.environ.get("RF_OFFLOAD_ARCH", "gfx950")
CLI_SOURCE = "cli/rockfit_cli.cpp"
CLI_OUT = os.path.join(HERE, "bin", "RockFit-DEM")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


# Flags of one source beyond the common ones.
SOURCE_FLAGS = {"csrc/rf_kernels.hip": ["-mllvm", "-amdgpu-sched-strategy=max-ilp"]}


def hipcc_cmd(src, obj, defines=()):
    """Compile one source to a position-independent object."""
    return (["hipcc", "-O3", f"--offload-arch={ARCH}", "-fPIC", "-c", "-std=c++17",
             "-ffp-contract=off", "-Wno-unused-result", "-Wno-unused-value",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include")] + SOURCE_FLAGS.get(src, []) +
            [f"-D{d}" for d in defines] + ["-o", obj, src])


def link_cmd(objs, out):
    return ["hipcc", f"--offload-arch={ARCH}", "-fPIC", "-shared", "-o", out] + list(objs)


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, cwd=HERE, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc build of librockfit_dem.so failed: " + " ".join(cmd))


def build(verbose=False, out=OUT, defines=()):
    """Compile the library; `defines` are tuning macros of the sources."""
    out = os.path.abspath(out)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    with tempfile.TemporaryDirectory(prefix="rockfit_build_") as tmp:
        objs = [os.path.join(tmp, os.path.basename(src) + ".o") for src in SOURCES]
        with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
            for f in [ex.submit(_run, hipcc_cmd(src, obj, defines), verbose)
                      for src, obj in zip(SOURCES, objs)]:
                f.result()
        _run(link_cmd(objs, out), verbose)
    return out


def cli_cmd(lib=OUT, out=CLI_OUT):
    libdir = os.path.dirname(os.path.abspath(lib))
    rel = os.path.relpath(libdir, os.path.dirname(os.path.abspath(out)))
    return ["g++", "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__",
            "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROCM, "include"),
            "-o", out, CLI_SOURCE, "-L" + libdir, "-lrockfit_dem",
            "-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
            "-Wl,-rpath,$ORIGIN/" + rel, "-Wl,-rpath," + os.path.join(ROCM, "lib")]


def build_cli(verbose=False, lib=OUT, out=CLI_OUT):
    """Compile bin/RockFit-DEM against the built library."""
    out = os.path.abspath(out)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    cmd = cli_cmd(lib, out)
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, cwd=HERE, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ build of RockFit-DEM failed")
    return out


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=OUT)
    ap.add_argument("-D", dest="defines", action="append", default=[])
    ap.add_argument("--no-cli", action="store_true")
    a = ap.parse_args()
    print(build(verbose=True, out=a.out, defines=a.defines))
    if not a.no_cli:
        print(build_cli(verbose=True, lib=a.out))

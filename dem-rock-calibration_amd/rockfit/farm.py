"""Driving the RockFit-DEM command line the way the reference's Python farm
does (NN_MP.py:71-215, 769-771): one process per sample, material in
``M(...)``, tests in the fourth argument, the index appended to ``I``, and
three stdout patterns read back.

The binary is bin/RockFit-DEM next to this package (built by build_lib.py).
"""
import os
import re
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINARY = os.path.join(HERE, "bin", "RockFit-DEM")

# NN_MP.py:769 -- ['./RockFit-DEM', 'G0', 'M(1,2,3,4,5,6)', '', 'I']
TEMPLATE = ("RockFit-DEM", "G0", "M(1,2,3,4,5,6)", "", "I")
# NN_MP.py:85-86, 771
INDEX_PATTERN = r"INDEX (\d+)"
GPU_PATTERN = r"Correct GPU request! Requested:\s+(\d+)"
RESULT_PATTERN = r"ResultCalculateOneSample\[(.*?)\]"

# ResultCalculateOneSample fields (mp_calculateonesample.cu:232-236)
RESULT_FIELDS = ("E", "sigma_c", "tau_c", "bond_radius", "gap", "porosity",
                 "strain_u", "stress_u", "linearity_deviation_u", "destruction_u",
                 "strain_b", "stress_b", "zero", "destruction_b")


def material_arg(values):
    """M(v0,...,v5) with each value as '%.8E' (NN_MP.py:317-320)."""
    if len(values) != 6:
        raise ValueError("six material values: E, sigma_c, tau_c, b_r, gap, porosity")
    return "M(" + ",".join("%.8E" % float(v) for v in values) + ")"


def command(values, tests="UB", gpu=0, index=0, binary=BINARY):
    """The argv of one farm task (template filled as NN_MP.py:92-93, 318-320)."""
    c = list(TEMPLATE)
    c[0] = binary
    c[1] = "G" + str(gpu)
    c[2] = material_arg(values)
    c[3] = tests
    c[4] = "I" + str(index)
    return c


def parse(stdout):
    """Index, freed GPU and result vector from one run's stdout; None where
    a pattern is absent."""
    out = {"index": None, "gpu": None, "result": None, "result_text": None}
    m = re.search(INDEX_PATTERN, stdout)
    if m:
        out["index"] = int(m.group(1))
    m = re.search(GPU_PATTERN, stdout)
    if m:
        out["gpu"] = int(m.group(1))
    m = re.search(RESULT_PATTERN, stdout)
    if m:
        text = m.group(1)
        out["result_text"] = text
        out["result"] = np.array(text.replace(" ", "").split(","), dtype=float)
    return out


def run(values, tests="UB", gpu=0, index=0, cwd=None, timeout=None, env=None, binary=BINARY):
    """Run one task; returns parse(stdout) plus 'stdout', 'stderr', 'returncode'."""
    if not os.path.exists(binary):
        raise FileNotFoundError(f"{binary} is not built (build_lib.py builds it)")
    p = subprocess.run(command(values, tests, gpu, index, binary), cwd=cwd, env=env,
                       capture_output=True, text=True, timeout=timeout)
    res = parse(p.stdout)
    res.update(stdout=p.stdout, stderr=p.stderr, returncode=p.returncode)
    return res


def append_result(path, result_text):
    """ResultMaterial.txt line: the captured group with commas as spaces
    (NN_MP.py:151-153)."""
    with open(path, "a") as f:
        f.write(result_text.replace(",", " "))
        f.write("\n")

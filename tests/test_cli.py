"""RockFit-DEM command line without a GPU: the argument grammar of
kernel.cu:23-124 (flags, the carried-over "CS"/"U"/"B" shift, INDEX echo),
the GPU-request lines the farm reads, and a loud failure when no device can
run the sample.

RF_CLI_DRY_RUN=1 makes the binary print its parsed flags
(G, M, CS, DCG, TSS, U, B, FUSC, FBSC in kernel.cu's l_flags order) and stop.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from rockfit import farm, sample

M = "M(1.71000000E+10,6.18000000E+08,5.03000000E+08,4.00000000E-04,5.00000000E-05,3.90000000E-01)"


@pytest.fixture(scope="module")
def binary():
    if not os.path.exists(farm.BINARY):
        import build_lib
        build_lib.build_cli()
    return farm.BINARY


def run(binary, args, dry=True, cwd=None):
    env = dict(os.environ)
    if dry:
        env["RF_CLI_DRY_RUN"] = "1"
    p = subprocess.run([binary] + list(args), capture_output=True, text=True, env=env, cwd=cwd,
                       timeout=120)
    return p


def flags_of(stderr):
    m = re.search(r"FLAGS ((?:[01] ?){9})", stderr)
    assert m, stderr
    return [int(v) for v in m.group(1).split()]


def gpu_count(stdout):
    m = re.match(r"There are (\d+) GPUs\n", stdout)
    assert m, stdout
    return int(m.group(1))


#                G  M  CS DCG TSS U  B  FUSC FBSC
CASES = [
    ([M, "UB", "I7"],                 [0, 1, 0, 0, 0, 1, 1, 0, 0]),
    ([M, "CSUB", "FUSC", "I12"],      [0, 1, 1, 0, 0, 1, 1, 1, 0]),
    ([M, "CS", "FBSC"],               [0, 1, 1, 0, 0, 0, 0, 0, 1]),
    # the shift set by CS stays: a later lone "U" or "B" is read at offset 2
    (["CS", "U", "B", M],             [0, 1, 1, 0, 0, 0, 0, 0, 0]),
    # "U" moves the offset to 1, so a following lone "B" is not seen
    (["U", "B"],                      [0, 0, 0, 0, 0, 1, 0, 0, 0]),
    (["B", "U"],                      [0, 0, 0, 0, 0, 0, 1, 0, 0]),
    (["UB", "FBSC", "DCG", "TSS"],    [0, 0, 0, 1, 1, 1, 1, 0, 1]),
    # the farm's empty fourth argument (NN_MP.py:769) matches nothing
    ([M, "", "I3"],                   [0, 1, 0, 0, 0, 0, 0, 0, 0]),
    # "CSB": U is tested at offset 2 ('B'), then B at offset 2
    (["CSB"],                         [0, 0, 1, 0, 0, 0, 1, 0, 0]),
]


@pytest.mark.parametrize("args,expect", CASES)
def test_argument_grammar(binary, args, expect):
    p = run(binary, args)
    assert p.returncode == 0, p.stderr
    assert flags_of(p.stderr) == expect
    assert "Selected by default GPU 0" in p.stderr


def test_index_echo_and_farm_patterns(binary):
    p = run(binary, [M, "UB", "I42"])
    lines = p.stdout.splitlines()
    assert lines[1] == "INDEX 42"
    got = farm.parse(p.stdout)
    assert got["index"] == 42 and got["gpu"] is None and got["result"] is None
    p = run(binary, ["I"])
    assert p.stdout.splitlines()[1] == "INDEX "
    assert farm.parse(p.stdout)["index"] is None


def test_material_string_is_kept_whole(binary):
    p = run(binary, [M])
    assert f"MATERIAL {M} " in p.stderr


def test_gpu_request_lines(binary):
    p = run(binary, ["G9", M, "UB", "I1"])
    n = gpu_count(p.stdout)
    assert n < 9
    assert p.returncode == 1
    assert "Incorrect GPU request! Requested:9" in p.stderr
    # kernel.cu returns at once: nothing after the device line
    assert p.stdout == f"There are {n} GPUs\n"
    p = run(binary, ["G0", M, "UB", "I1"])
    if n == 0:
        assert p.returncode == 1 and "Incorrect GPU request! Requested:0" in p.stderr
    else:
        assert p.returncode == 0
        assert p.stdout.splitlines()[1] == "Correct GPU request! Requested: 0"
        assert farm.parse(p.stdout)["gpu"] == 0


def test_farm_command_matches_template():
    c = farm.command([1.71e10, 6.18e8, 5.03e8, 4e-4, 5e-5, 0.39], "UB", gpu=2, index=17,
                     binary="./RockFit-DEM")
    assert c == ["./RockFit-DEM", "G2", M, "UB", "I17"]


def test_farm_parse_result_line():
    text = ("There are 1 GPUs\nCorrect GPU request! Requested: 0\nINDEX 5\n"
            "ResultCalculateOneSample[1.71e+10,6.18e+08,5.03e+08,0.0004,5e-05,0.39,"
            "0.00480002,1.04e+08,1.2,0.9,0.01,1e+06,0,1]\nTime x | y\n")
    got = farm.parse(text)
    assert got["index"] == 5 and got["gpu"] == 0
    assert got["result"].shape == (14,)
    assert got["result"][1] == 6.18e8 and got["result"][12] == 0.0


def test_fails_loudly_without_a_device(binary, tmp_path):
    if gpu_count(run(binary, []).stdout) > 0:
        pytest.skip("a GPU is present")
    os.makedirs(tmp_path / "result")
    xyz = np.array([[20.0, 20.0, 30.0], [21.0, 20.0, 30.0]])
    sample.write_sample(str(tmp_path / "result" / "Sample_0_0.txt"), xyz)
    p = run(binary, [M, "U", "I1"], dry=False, cwd=str(tmp_path))
    assert p.returncode != 0
    assert "rf_create failed" in p.stderr
    assert "ResultCalculateOneSample" not in p.stdout


def test_missing_sample_file_is_an_error(binary, tmp_path):
    p = run(binary, [M, "B", "I1"], dry=False, cwd=str(tmp_path))
    assert p.returncode != 0
    assert "Sample_1_0.txt" in p.stderr

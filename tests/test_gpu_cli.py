"""RockFit-DEM on the GPU: the command line's ResultCalculateOneSample line
and curve files equal what the Python API computes through the same C-ABI
on the same sample file (which the loading parity tests pin to the oracle).

Small specimens (1500 spheres of radius 0.5 mm, the reference's p_r) are
written as ./result/Sample_<k>_0.txt in a temporary working directory and
placed at the centre of the reference's space boxes (40x40x60 mm uniaxial,
65x65x30 mm Brazilian; mp_calculateonesample.cu:91).
"""
import os

import numpy as np
import pytest

from conftest import gpu_available
from rockfit import config, core, farm, sample
from rockfit.core import Context

pytestmark = pytest.mark.gpu

MAT = [1.71e10, 6.18e8, 5.03e8, 4.0e-4, 5.0e-5, 0.39]
SPACE = ([40.0, 40.0, 60.0], [65.0, 65.0, 30.0])
P_R = 0.5


@pytest.fixture(autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.skip("no GPU")


@pytest.fixture(scope="module")
def binary():
    assert os.path.exists(farm.BINARY), "bin/RockFit-DEM is not built"
    return farm.BINARY


def make_sample(kind, n, seed):
    shape, aspect = ("cylinder", 2.0) if kind == 0 else ("disc", 0.5)
    xyz, _ = sample.generate(n, P_R, P_R, shape=shape, aspect=aspect, seed=seed, relax="gpu")
    box = core.box_reference(SPACE[kind], 2.0 * 3.2 * P_R)
    return xyz + np.array(box.center[:])


def api_run(kind, path, values):
    """The same run through the Python API: h_LoadCylinderSample's size_real
    and center_real (md_createcylindersample.cu:448-465), driver values of
    mp_calculateonesample.cu:120-124."""
    xyz, _ = sample.read_sample(path)
    r = np.full(len(xyz), P_R)
    mat = config.reference_material(E=values[0], sigma_c=values[1], tau_c=values[2],
                                    bond_radius_m=values[3], gap_m=values[4])
    box = core.box_reference(SPACE[kind], 2.0 * mat.acut)
    if kind == 0:
        cfg = config.uniaxial_cfg(xyz, r, v_load_mps=-0.1, dnsave=100)
    else:
        cfg = config.brazilian_cfg(xyz, r, v_load_mps=-0.1, dnsave=1000)
    lo, hi = xyz.min(axis=0), xyz.max(axis=0)
    for k in range(3):
        cfg.size[k] = hi[k] - lo[k] + 2.0 * P_R
        cfg.center[k] = 0.5 * (hi[k] + lo[k])
    g = Context(xyz, r, mat, box)
    g.build_bonds()
    g.set_test(cfg)
    ts = g.run()
    curve = g.curve()
    g.close()
    return mat, ts, curve


def fmt(v):
    return "%g" % v


def check_curve_file(path, curve):
    lines = open(path).read().splitlines()
    assert lines[0] == "EpsZ StressZ_bottom StressZ_top"
    assert len(lines) == len(curve) + 1
    for line, row in zip(lines[1:], curve):
        assert line == " ".join([fmt(row[0]), fmt(row[1] * 1e9), fmt(row[2] * 1e9)])


def test_cli_uniaxial_equals_api(binary, tmp_path):
    print("case cli uniaxial n=1500 seed=5", flush=True)
    os.makedirs(tmp_path / "result")
    path = str(tmp_path / "result" / "Sample_0_0.txt")
    sample.write_sample(path, make_sample(0, 1500, 5))
    res = farm.run(MAT, "U", gpu=0, index=5, cwd=str(tmp_path), timeout=300)
    assert res["returncode"] == 0, res["stderr"][-2000:]
    assert res["index"] == 5 and res["gpu"] == 0
    mat, ts, curve = api_run(0, path, MAT)
    assert ts.loading, "the platens never loaded the specimen"
    want = [mat.E_bond * 1e9, mat.sigma_c * 1e9, mat.tau_c * 1e9, mat.bond_radius * 1e-3,
            mat.bond_gap * 1e-3, MAT[5], ts.peak_strain, ts.peak_stress * 1e9,
            ts.linearity_deviation, ts.destruction, 1e-2, 1e6, 0, 1.0]
    assert res["result_text"] == ",".join(fmt(v) for v in want)
    check_curve_file(str(tmp_path / "result" / "Uniaxial_0.txt"), curve)


def test_cli_brazilian_equals_api(binary, tmp_path):
    print("case cli brazilian n=1500 seed=6", flush=True)
    os.makedirs(tmp_path / "result")
    path = str(tmp_path / "result" / "Sample_1_0.txt")
    sample.write_sample(path, make_sample(1, 1500, 6))
    res = farm.run(MAT, "B", gpu=0, index=6, cwd=str(tmp_path), timeout=300)
    assert res["returncode"] == 0, res["stderr"][-2000:]
    mat, ts, curve = api_run(1, path, MAT)
    assert ts.loading
    want = [mat.E_bond * 1e9, mat.sigma_c * 1e9, mat.tau_c * 1e9, mat.bond_radius * 1e-3,
            mat.bond_gap * 1e-3, MAT[5], 1e-2, 1e6, 1.0, 1.0, ts.peak_strain,
            ts.peak_stress * 1e9, 0, ts.destruction]
    assert res["result_text"] == ",".join(fmt(v) for v in want)
    check_curve_file(str(tmp_path / "result" / "Brazil_0.txt"), curve)

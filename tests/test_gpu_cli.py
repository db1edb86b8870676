"""RockFit-DEM on the GPU: the command line's ResultCalculateOneSample line
and curve files equal what the Python API computes through the same C-ABI
on the same sample file, and what the CPU oracle computes on that file
under the reference's configuration (SetInitialParameters,
mp_determineconjgradient.cu:15-47: p_r 0.5 mm, one modulus for bonds and
contacts, acut 3.2 p_r, the DCG runs re-gridded at 2.9 p_r (:65-66), the
reference's quirks).

Small specimens (1500 spheres of radius 0.5 mm, the reference's p_r) are
written as ./result/Sample_<k>_0.txt in a temporary working directory and
placed at the centre of the reference's space boxes (40x40x60 mm uniaxial,
65x65x30 mm Brazilian; mp_calculateonesample.cu:91).
"""
import os
import subprocess

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


def api_run(kind, path, values, oracle=False):
    """The same run through the Python API: h_LoadCylinderSample's size_real
    and center_real (md_createcylindersample.cu:448-465), driver values of
    mp_calculateonesample.cu:120-124.  oracle: also run it in the CPU
    oracle and return its test state and curve."""
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
    if oracle:
        return mat, ts, curve, oracle_run(kind, path, mat, box, cfg)
    return mat, ts, curve


def fmt(v):
    return "%g" % v


def oracle_run(kind, path, mat, box, cfg):
    """The whole test in the CPU oracle on the same sample file: lists,
    bonds, loading to the end, CalculateLinearityDeviation and
    CalculateDestruction."""
    import pyoracle

    xyz, _ = sample.read_sample(path)
    r = np.full(len(xyz), P_R)
    o = pyoracle.Oracle(xyz, r, mat, box, threads=16)
    o.build_bonds()
    o.set_test(cfg)
    o.step(1 << 40)
    o.post()
    ts = o.test_state()
    assert ts.done
    return ts, o.curve()


def check_curve_file(path, curve):
    lines = open(path).read().splitlines()
    assert lines[0] == "EpsZ StressZ_bottom StressZ_top"
    assert len(lines) == len(curve) + 1
    for line, row in zip(lines[1:], curve):
        assert line == " ".join([fmt(row[0]), fmt(row[1] * 1e9), fmt(row[2] * 1e9)])


@pytest.mark.timeout(900)
def test_cli_uniaxial_equals_api_and_oracle(binary, tmp_path):
    print("case cli uniaxial n=1500 seed=5", flush=True)
    os.makedirs(tmp_path / "result")
    path = str(tmp_path / "result" / "Sample_0_0.txt")
    sample.write_sample(path, make_sample(0, 1500, 5))
    res = farm.run(MAT, "U", gpu=0, index=5, cwd=str(tmp_path), timeout=300)
    assert res["returncode"] == 0, res["stderr"][-2000:]
    assert res["index"] == 5 and res["gpu"] == 0
    mat, ts, curve, (ots, ocurve) = api_run(0, path, MAT, oracle=True)
    assert ts.loading, "the platens never loaded the specimen"
    for t in (ts, ots):
        want = [mat.E_bond * 1e9, mat.sigma_c * 1e9, mat.tau_c * 1e9, mat.bond_radius * 1e-3,
                mat.bond_gap * 1e-3, MAT[5], t.peak_strain, t.peak_stress * 1e9,
                t.linearity_deviation, t.destruction, 1e-2, 1e6, 0, 1.0]
        assert res["result_text"] == ",".join(fmt(v) for v in want)
    print(f"uniaxial: {ots.step} loading steps, UCS {ots.peak_stress * 1e3:.4f} MPa, "
          f"equal to the oracle", flush=True)
    np.testing.assert_array_equal(curve, ocurve)
    check_curve_file(str(tmp_path / "result" / "Uniaxial_0.txt"), ocurve)


@pytest.mark.timeout(900)
def test_cli_brazilian_equals_api_and_oracle(binary, tmp_path):
    print("case cli brazilian n=1500 seed=6", flush=True)
    os.makedirs(tmp_path / "result")
    path = str(tmp_path / "result" / "Sample_1_0.txt")
    sample.write_sample(path, make_sample(1, 1500, 6))
    res = farm.run(MAT, "B", gpu=0, index=6, cwd=str(tmp_path), timeout=300)
    assert res["returncode"] == 0, res["stderr"][-2000:]
    mat, ts, curve, (ots, ocurve) = api_run(1, path, MAT, oracle=True)
    assert ts.loading
    for t in (ts, ots):
        want = [mat.E_bond * 1e9, mat.sigma_c * 1e9, mat.tau_c * 1e9, mat.bond_radius * 1e-3,
                mat.bond_gap * 1e-3, MAT[5], 1e-2, 1e6, 1.0, 1.0, t.peak_strain,
                t.peak_stress * 1e9, 0, t.destruction]
        assert res["result_text"] == ",".join(fmt(v) for v in want)
    print(f"brazilian: {ots.step} loading steps, tensile strength {ots.peak_stress * 1e3:.4f} "
          f"MPa, equal to the oracle", flush=True)
    np.testing.assert_array_equal(curve, ocurve)
    check_curve_file(str(tmp_path / "result" / "Brazil_0.txt"), ocurve)


def cg_box(kind):
    """DetermineConjGradient's run box: space box set up with cell 2 x 3.2 p_r,
    re-gridded with 2 x 2.9 p_r."""
    b = core.box_reference(SPACE[kind], 2.0 * 3.2 * P_R)
    cell = 2.0 * 2.9 * P_R
    b.cell = cell
    for d in range(3):
        B = b.center[d] + b.half[d]
        b.dims[d] = int(np.floor(B * (1.0 / cell))) + (4 if d == 2 else 1)
    return b


def cg_material(P):
    m = config.reference_material()
    m.E_bond = m.E_contact = P[0]
    m.sigma_c, m.tau_c, m.bond_radius, m.bond_gap = P[1], P[2], P[3], P[4]
    m.acut = 2.9 * P_R
    return m


@pytest.mark.timeout(900)
def test_cli_dcg_step_uses_the_right_runs(binary, tmp_path):
    """Two steps of DetermineConjGradient on a small uniaxial sample: the
    base run and the three varied runs (type 7: sigma_c, tau_c, b_r, each
    +10 %) go to the GPU as one batch; the recorded results and functional
    equal single runs through the API, and the next point is the CG update
    of those results (goals and weights of mp_determineconjgradient.cu:150-160,
    273-276)."""
    print("case cli DCG n=1500 seed=9 steps=2", flush=True)
    os.makedirs(tmp_path / "result")
    path = str(tmp_path / "result" / "Sample_0.txt")
    xyz0, _ = sample.generate(1500, P_R, P_R, shape="cylinder", aspect=2.0, seed=9, relax="gpu")
    cbox = core.box_reference(SPACE[0], 2.0 * 3.2 * P_R)
    sample.write_sample(path, xyz0 + np.array(cbox.center[:]))
    env = dict(os.environ, RF_CG_STEPS="2", RF_CG_REUSE_SAMPLES="1")
    p = subprocess.run([binary, "G0", "DCG"], cwd=str(tmp_path), env=env, capture_output=True,
                       text=True, timeout=600)
    assert p.returncode == 0, p.stderr[-3000:]
    lines = open(tmp_path / "result" / "CalculationResults.dat").read().splitlines()
    assert len(lines) == 8
    xyz, _ = sample.read_sample(path)
    r = np.full(len(xyz), P_R)
    lo, hi = xyz.min(axis=0), xyz.max(axis=0)
    cfg = config.uniaxial_cfg(xyz, r, v_load_mps=-0.1, min_stress_mpa=10.0, max_strain=0.025,
                              cond_fracture=0.2, dnsave=100)
    for k in range(3):
        cfg.size[k] = hi[k] - lo[k] + 2.0 * P_R
        cfg.center[k] = 0.5 * (hi[k] + lo[k])
    box = cg_box(0)

    def results(P, oracle=False):
        g = Context(xyz, r, cg_material(P), box)
        g.build_bonds()
        g.set_test(cfg)
        ts = g.run()
        g.close()
        out = np.array([ts.peak_strain, ts.peak_stress, 1.0, ts.linearity_deviation])
        if oracle:  # the base run of the first step, in the CPU oracle
            ots, _ = oracle_run(0, path, cg_material(P), box, cfg)
            np.testing.assert_array_equal(
                out, [ots.peak_strain, ots.peak_stress, 1.0, ots.linearity_deviation])
        return out

    goal = np.array([0.00480002, 104.0497482e6 / 1e9, 10.5e6 / 1e9, 1.0])
    W = np.array([50.0, 100.0, 1.0, 10000.0]) / goal ** 2
    W[2], goal[2] = 100.0, 1.0
    varied = np.array([0, 1, 1, 1, 0])
    P = np.array([0.45 * 1.71066e10 / 1e9, 1.0 * 6.18389e8 / 1e9, 1.0 * 5.03332e8 / 1e9,
                  4 * 0.000409345 / 1e-3, 0.00046797 / 1e-3])
    dF_old = np.zeros(5)
    for step in range(2):
        base = lines[4 * step:4 * step + 4]
        assert base[0] == f"Step: {step}"
        got = np.array([float(v) for v in base[1].split(":")[1].split()])
        np.testing.assert_allclose(got, P * np.array([1e9, 1e9, 1e9, 1e-3, 1e-3]), rtol=1e-5)
        var = varied * 0.1 * P
        R0 = results(P, oracle=step == 0)
        assert R0[1] > 0.0
        want = "%g %g %g" % (R0[0], R0[1] * 1e9, R0[2] * 1e9)
        assert base[2].split(": ")[1] == want
        dR = np.zeros((5, 4))
        for i in range(5):
            if var[i] > 1e-12:
                Pi = P.copy()
                Pi[i] += var[i]
                dR[i] = (results(Pi) - R0) / var[i]
        D = goal - R0
        F = W * D * D
        got_F = np.array([float(v) for v in base[3].split()[1:5]])
        np.testing.assert_allclose(got_F, F, rtol=1e-5, atol=1e-12)
        if step == 1:
            break  # the update after the last step is not recorded
        g = -2.0 * dR @ (W * D)
        beta = (g @ g) / (dF_old @ dF_old) if dF_old @ dF_old > 1e-12 else 0.0
        dd = g + beta * dF_old
        s = dR.T @ dd
        t = -np.sum(W * D * s) / np.sum(W * s * s)
        P = P - t * dd
        P = np.maximum(P, [1e5 / 1e9, 1e5 / 1e9, 1e8 / 1e9, 1e-6 / 1e-3, 1e-5 / 1e-3])
        dF_old = g

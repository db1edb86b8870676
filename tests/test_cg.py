"""The command line's conjugate-gradient calibration step (cli/rf_cg.h, the
reference's mp_task_data CG of mp_task.h:134-321) on linear response models
R = A P + b, where forward differences are exact and the mathematics fixes
the answers:

- the first step is an exact line search along the gradient of
  F = sum_j W_j (G_j - R_j)^2 (steepest descent; textbook step length
  g'g / (2 g'A'WA g));
- every step is an exact line search along its direction, so F never
  increases and the new gradient is orthogonal to the last direction;
- the direction is g_k + beta g_(k-1), beta = |g_k|^2 / |g_(k-1)|^2 (the
  reference overwrites its direction with the old gradient first);
- only the parameters the variation type moves change; SetMaterialContractions
  holds the lower bounds; SaveResults writes the reference's four lines.
"""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "dem-rock-calibration_amd", "cli")
HERE = os.path.dirname(os.path.abspath(__file__))

VARIED = [[1, 1, 1, 0, 0], [0, 0, 0, 1, 1], [1, 1, 1, 1, 1], [0, 1, 1, 1, 1], [0, 0, 1, 1, 1],
          [1, 0, 0, 1, 1], [1, 0, 0, 0, 0], [0, 1, 1, 1, 0]]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("cg") / "cg_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I" + CLI, "-o", exe,
                    os.path.join(HERE, "cg", "cg_check.cpp")], check=True)
    return exe


def run(checker, tmp_path, A, b, P0, goal, mult, vtype=2, delta=0.1, steps=1, u=1, bt=1):
    def fmt(v):
        return " ".join("%.17g" % float(x) for x in v)

    lines = [f"{vtype} {delta!r} {steps} {u} {bt}", fmt(goal), fmt(mult), fmt(P0)]
    lines += [fmt(row) for row in A] + [fmt(b)]
    out_file = str(tmp_path / "CalculationResults.dat")
    if os.path.exists(out_file):
        os.remove(out_file)
    p = subprocess.run([checker, out_file], input="\n".join(lines) + "\n", capture_output=True,
                       text=True, check=True)
    steps_out, runs = [], []
    cur = None
    for line in p.stdout.splitlines():
        w = line.split()
        if w[0] == "weights":
            weights = np.array(w[1:], float)
        elif w[0] == "run":
            runs.append((int(w[1]), int(w[2])))
        elif w[0] == "step":
            cur = {"t": float(w[3])}
            steps_out.append(cur)
        else:
            cur[w[0]] = np.array(w[1:], float)
    return weights, steps_out, runs, open(out_file).read()


def model(seed=1):
    rng = np.random.default_rng(seed)
    # well-conditioned response of 4 results to 5 parameters of order 1
    A = rng.normal(0.0, 1.0, (4, 5)) + 3.0 * np.eye(4, 5)
    b = rng.normal(0.0, 0.5, 4)
    P0 = rng.uniform(1.0, 2.0, 5)
    return A, b, P0


def functional(A, b, P, goal, W):
    D = goal - (A @ P + b)
    return float(np.sum(W * D * D)), D


def test_weights_and_first_step_is_exact_line_search(checker, tmp_path):
    A, b, P0 = model(1)
    goal = np.array([3.0, 4.0, 5.0, 6.0])
    mult = np.array([50.0, 100.0, 1.0, 10000.0])
    W, st, _, _ = run(checker, tmp_path, A, b, P0, goal, mult, vtype=2)
    np.testing.assert_allclose(W, mult / goal ** 2, rtol=1e-15)
    _, D = functional(A, b, P0, goal, W)
    g = -2.0 * A.T @ (W * D)
    tau = (g @ g) / (2.0 * g @ (A.T @ (W[:, None] * A)) @ g)
    np.testing.assert_allclose(st[0]["dF"], g, rtol=1e-9, atol=1e-12 * np.abs(g).max())
    np.testing.assert_allclose(st[0]["dd"], g, rtol=1e-9, atol=1e-12 * np.abs(g).max())
    np.testing.assert_allclose(st[0]["params"], P0 - tau * g, rtol=1e-9)
    F0, _ = functional(A, b, P0, goal, W)
    np.testing.assert_allclose(st[0]["F"].sum(), F0, rtol=1e-12)


def test_line_searches_decrease_and_direction_rule(checker, tmp_path):
    A, b, P0 = model(2)
    goal = np.array([3.0, 4.0, 5.0, 6.0])
    mult = np.array([50.0, 100.0, 1.0, 10000.0])
    W, st, _, _ = run(checker, tmp_path, A, b, P0, goal, mult, vtype=2, steps=25)
    P = [P0] + [s["params"] for s in st]
    F = [functional(A, b, p, goal, W)[0] for p in P]
    assert all(F[k + 1] <= F[k] * (1 + 1e-12) for k in range(len(F) - 1))
    # the reference's direction (g_k + beta g_(k-1)) is not the conjugate
    # one, so the decrease is steady but slower than n steps of linear CG
    assert F[-1] < 1e-3 * F[0]
    for k in range(1, len(st)):
        g_old, g_new = st[k - 1]["dF"], st[k]["dF"]
        beta = (g_new @ g_new) / (g_old @ g_old)
        np.testing.assert_allclose(st[k]["dd"], g_new + beta * g_old, rtol=1e-9,
                                   atol=1e-9 * np.abs(g_new).max())
        # the gradient at the new point is orthogonal to the last direction
        _, D = functional(A, b, P[k], goal, W)
        g_here = -2.0 * A.T @ (W * D)
        d = st[k - 1]["dd"]
        assert abs(g_here @ d) <= 1e-7 * np.linalg.norm(g_here) * np.linalg.norm(d) + 1e-14


def test_variation_type_moves_only_its_parameters(checker, tmp_path):
    A, b, P0 = model(3)
    goal = np.array([3.0, 4.0, 5.0, 6.0])
    mult = np.array([50.0, 100.0, 1.0, 10000.0])
    for vtype in (7, 1, 6):
        _, st, runs, _ = run(checker, tmp_path, A, b, P0, goal, mult, vtype=vtype, steps=3)
        fixed = [i for i in range(5) if VARIED[vtype][i] == 0]
        for s in st:
            np.testing.assert_array_equal(s["params"][fixed], P0[fixed])
        # run k > 0 is made only when parameter k-1 varies
        for run_k, made in runs[:6]:
            assert made == (1 if run_k == 0 else VARIED[vtype][run_k - 1])


def test_lower_bounds(checker, tmp_path):
    """A target far below the start drives every parameter down: each stops
    at its SetMaterialContractions bound (reduced units)."""
    A = np.eye(4, 5) * 1.0
    A[0, 4] = 1.0
    b = np.zeros(4)
    P0 = np.array([1.0, 1.0, 1.0, 1.0, 1.0])
    goal = np.array([-50.0, -50.0, -50.0, -50.0])
    mult = np.array([1.0, 1.0, 1.0, 1.0])
    _, st, _, _ = run(checker, tmp_path, A, b, P0, goal, mult, vtype=2)
    bounds = np.array([1e5 / 1e9, 1e5 / 1e9, 1e8 / 1e9, 1e-6 / 1e-3, 1e-5 / 1e-3])
    np.testing.assert_array_equal(st[0]["params"], bounds)


def test_tests_off_weights(checker, tmp_path):
    """Without the uniaxial test the reference sets weights 0 and 3 (and
    goals 0, 1, 3) to 1 but leaves weight 1 as computed; without the
    Brazilian test weight 2 is 100 and goal 2 is 1."""
    A, b, P0 = model(4)
    goal = np.array([3.0, 4.0, 5.0, 6.0])
    mult = np.array([50.0, 100.0, 1.0, 10000.0])
    W, _, _, _ = run(checker, tmp_path, A, b, P0, goal, mult, u=0, bt=0)
    np.testing.assert_allclose(W, [1.0, 100.0 / 16.0, 100.0, 1.0], rtol=1e-15)


def test_save_results_format(checker, tmp_path):
    A, b, P0 = model(5)
    P0 = P0 * np.array([1e-1, 1e-3, 1e-3, 1.0, 1.0])
    goal = np.array([3.0, 4.0, 5.0, 6.0])
    mult = np.array([50.0, 100.0, 1.0, 10000.0])
    W, st, _, text = run(checker, tmp_path, A, b, P0, goal, mult, vtype=2, steps=2)
    lines = text.splitlines()
    assert len(lines) == 8
    assert lines[0] == "Step: 0" and lines[4] == "Step: 1"
    head = "Parameters (E, Ecrit, Gcrit, boundR, boundDistance): "
    assert lines[1] == head + " ".join("%g" % v for v in
                                       [P0[0] * 1e9, P0[1] * 1e9, P0[2] * 1e9, P0[3] * 1e-3,
                                        P0[4] * 1e-3]) + " "
    R0 = A @ P0 + b
    assert lines[2] == ("Results (UniaxialFractureStrain UniaxialFractureStress "
                        "BrasilFractureStress): %g %g %g" % (R0[0], R0[1] * 1e9, R0[2] * 1e9))
    m = re.match(r"DFunctional (\S+) (\S+) (\S+) (\S+) \| (\S+)$", lines[3])
    assert m
    vals = np.array([float(v) for v in m.groups()])
    np.testing.assert_allclose(vals[:4], st[0]["F"], rtol=1e-5)
    np.testing.assert_allclose(vals[4], st[0]["F"].sum(), rtol=1e-5)


def test_python_cg_equals_the_command_line_step(checker, tmp_path):
    """rockfit.calibrate.ConjugateGradient (the CG step with the Jacobian
    given) lands where cli/rf_cg.h's step lands on a linear model, whose
    forward differences are exact: three steps, all five parameters."""
    from rockfit.calibrate import ConjugateGradient

    A, b, P0 = model(6)
    goal = np.array([3.0, 4.0, 5.0, 6.0])
    mult = np.array([50.0, 100.0, 1.0, 10000.0])
    W, st, _, _ = run(checker, tmp_path, A, b, P0, goal, mult, vtype=2, steps=3)
    cg = ConjugateGradient(goal, mult)
    np.testing.assert_array_equal(cg.weight, W)
    P = P0.copy()
    for k in range(3):
        P = cg.step(P, A @ P + b, A.T)
        np.testing.assert_allclose(P, st[k]["params"], rtol=1e-8)
        np.testing.assert_allclose(cg.dF, st[k]["dF"], rtol=1e-7, atol=1e-9 * np.abs(cg.dF).max())
        np.testing.assert_allclose(cg.F, st[k]["F"], rtol=1e-8)

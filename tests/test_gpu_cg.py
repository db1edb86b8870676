"""The reference's conjugate-gradient step (mp_task.h:233-311) driven by
the tangents of one GPU run instead of NParam + 1 forward-difference runs.

Response: the mean axial stress and the secant modulus of an elastic
loading window (2000 spheres, 300 steps); parameters: bond E and bond radius
ratio; goals: the response at (1.2 E, 0.9 ratio).

Bar: on the frozen-branch model (the base window's decisions replayed), the
step from the tangents and the step from forward differences at relative
step 1e-6 (the reference's Jacobian, mp_determineconjgradient.cu:84-142)
land on the same next point to 1e-5 relative, with 1 run against 3; and
three gradient steps on the plain model lower the functional.
"""
import numpy as np
import pytest

from rockfit import autograd, calibrate, config

pytestmark = pytest.mark.gpu

NAMES = ["E_bond", "bond_radius_ratio"]


@pytest.fixture(scope="module")
def setup():
    w = config.gradient_workload(n=2000, steps=300, min_stress_mpa=1e4)
    return dict(xyz=w["xyz"], r=w["r"], mat=w["mat"], box=w["box"], cfg=w["cfg"], steps=300)


def _theta(setup):
    return np.array([getattr(setup["mat"], autograd.PARAMS[n][1]) for n in NAMES])


def test_gradient_step_equals_difference_step(setup):
    print("case CG step, tangents vs forward differences, n=2000", flush=True)
    th0 = _theta(setup)
    goal, _ = calibrate.window_response(setup, NAMES, th0 * np.array([1.2, 0.9]))
    s0, _, e0, rec = autograd.run_frozen(setup, NAMES, th0)
    sa = 0.5 * (np.abs(s0[:, 0]) + np.abs(s0[:, 1]))
    frozen = (rec, int(np.argmax(sa)))
    mult = np.array([1.0, 1.0])
    cg_t = calibrate.ConjugateGradient(goal, mult)
    th_t, R_t = calibrate.cg_step_gradient(cg_t, setup, NAMES, th0, frozen=frozen)
    cg_d = calibrate.ConjugateGradient(goal, mult)
    th_d, R_d = calibrate.cg_step_differences(cg_d, setup, NAMES, th0, 1e-6, frozen=frozen)
    np.testing.assert_array_equal(R_t, R_d)
    rel = np.abs(th_t - th_d) / np.abs(th_t)
    print(f"next point from tangents {th_t}, from differences {th_d}, rel {rel}", flush=True)
    assert np.all(rel < 1e-5), rel
    # the step moves the point (mostly along the ratio: both responses
    # follow the bond stiffness, which goes as E ratio^2)
    assert np.max(np.abs(th_t - th0) / np.abs(th0)) > 1e-2
    # the reference's own 10 % forward differences, for scale (not asserted)
    cg_r = calibrate.ConjugateGradient(goal, mult)
    th_r, _ = calibrate.cg_step_differences(cg_r, setup, NAMES, th0, 0.1)
    print(f"next point from 10 % differences (plain model) {th_r}", flush=True)


def test_gradient_cg_lowers_the_functional(setup):
    print("case gradient CG iterations, n=2000", flush=True)
    th = _theta(setup)
    goal, _ = calibrate.window_response(setup, NAMES, th * np.array([1.2, 0.9]))
    cg = calibrate.ConjugateGradient(goal, np.array([1.0, 1.0]))
    F = []
    for k in range(3):
        th, R = calibrate.cg_step_gradient(cg, setup, NAMES, th)
        F.append(float(cg.F.sum()))
        print(f"step {k}: F {F[-1]:.4e} next {th}", flush=True)
    R, _ = calibrate.window_response(setup, NAMES, th)
    D = goal - R
    F.append(float(np.sum(cg.weight * D * D)))
    print(f"final F {F[-1]:.4e}", flush=True)
    assert F[-1] < 0.05 * F[0], F

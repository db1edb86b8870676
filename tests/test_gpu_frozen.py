"""Tangents against central differences of the frozen-branch model.

A sensitivity window records every discrete decision it takes (bond breaks,
contact touch and slip, rolling friction and its chatter regime with the
spin held fixed, platen contacts and slip, the controller's platen velocity:
rf_sens_branches).  Windows at nearby parameters replay those decisions, so
their central differences converge to the derivative the tangents carry.

Bar: the replay at the recorded parameters repeats the recorded curve bit
for bit; at relative step 1e-6 the central differences of the curve
stresses (2000 spheres, 300 steps, every curve point) and of the mean axial
stress and secant modulus (the 20k-sphere, 2000-step window to 0.1 %
strain, peak held at the recorded point) agree with the tangents to 1e-5
relative, for bond E, bond radius ratio and contact E.
"""
import numpy as np
import pytest

from rockfit import autograd, config

pytestmark = pytest.mark.gpu

NAMES = ["E_bond", "bond_radius_ratio", "E_contact"]
REL = 1e-6
TOL = 1e-5


def _setup(n, steps, min_stress_mpa=0.1):
    w = config.gradient_workload(n=n, steps=steps, min_stress_mpa=min_stress_mpa)
    return dict(xyz=w["xyz"], r=w["r"], mat=w["mat"], box=w["box"], cfg=w["cfg"], steps=steps)


def _theta(setup):
    return np.array([getattr(setup["mat"], autograd.PARAMS[nm][1]) for nm in NAMES])


def _replays(setup, rec, k, rel):
    th = _theta(setup)
    h = rel * abs(th[k])
    tp, tm = th.copy(), th.copy()
    tp[k] += h
    tm[k] -= h
    sp, _, ep, cp = autograd.run_frozen(setup, NAMES, tp, recorded=rec)
    sm, _, em, cm = autograd.run_frozen(setup, NAMES, tm, recorded=rec)
    cp.close()
    cm.close()
    return sp, ep, sm, em, h


def test_frozen_replay_and_curve_tangents():
    # a fast window: the loading threshold (and with it the fracture test)
    # is set out of reach, so the test does not end at the first impact
    setup = _setup(2000, 300, min_stress_mpa=1e4)
    print("case frozen branches n=2000 steps=300", flush=True)
    th = _theta(setup)
    s0, J, e0, rec = autograd.run_frozen(setup, NAMES, th)
    s1, _, e1, c1 = autograd.run_frozen(setup, NAMES, th, recorded=rec)
    np.testing.assert_array_equal(s1, s0)
    np.testing.assert_array_equal(e1, e0)
    c1.close()
    assert np.any(s0 != 0.0)
    for k, name in enumerate(NAMES):
        sp, _, sm, _, h = _replays(setup, rec, k, REL)
        fd = (sp - sm) / (2 * h)
        err = np.max(np.abs(J[:, k, :] - fd)) / np.max(np.abs(J[:, k, :]))
        print(f"{name}: curve tangents vs frozen central differences (rel {REL:g}): {err:.2e}",
              flush=True)
        assert err < TOL, (name, err)


@pytest.mark.timeout(600)
def test_frozen_macro_gradient_20k():
    setup = _setup(20000, 2000)
    print("case frozen branches n=20000 steps=2000", flush=True)
    th = _theta(setup)
    s0, J, e0, rec = autograd.run_frozen(setup, NAMES, th)
    sa = 0.5 * (np.abs(s0[:, 0]) + np.abs(s0[:, 1]))
    peak = int(np.argmax(sa))
    v0, jac = autograd.macro_from_curve(e0, s0, J, peak=peak)
    assert v0[0] > 0.0 and v0[1] > 0.0
    for k, name in enumerate(NAMES):
        sp, ep, sm, em, h = _replays(setup, rec, k, REL)
        fd = (autograd.macro_from_curve(ep, sp, peak=peak)[0]
              - autograd.macro_from_curve(em, sm, peak=peak)[0]) / (2 * h)
        err = np.abs(jac[:, k] - fd) / np.abs(jac[:, k])
        print(f"{name}: mean stress tangent {jac[0, k]:.6e} fd {fd[0]:.6e}; secant modulus "
              f"tangent {jac[1, k]:.6e} fd {fd[1]:.6e}; rel err {err[0]:.2e} {err[1]:.2e}",
              flush=True)
        assert np.all(err < TOL), (name, err)

"""GPU sensitivities (rf_sens_*, rockfit.autograd) against central finite
differences of the plain GPU step.

Loading window of a bonded 2000-sphere sample: the platens close in, touch
and load the specimen to about 3 MPa before the first bond breaks.

What finite differences can check: on a short window (30 steps) central
differences converge for relative steps 1e-4 ... 1e-6, and are the
reference; on longer windows the stick/slip and rolling-friction switches of
the contact law make small-step differences erratic (they scatter by orders
of magnitude between 1e-5 and 1e-7), while a 3 % step gives the smooth
response's slope.  The tangents are exact except for the direction of a
rolling-friction torque in its chatter regime (rf_physics.h), which costs
about 1 %.
"""
import numpy as np
import pytest
import torch

from conftest import gpu_available
from rockfit import autograd, config
from rockfit.core import Context

pytestmark = pytest.mark.gpu

NAMES = ["E_bond", "nu", "E_contact", "friction"]


@pytest.fixture(autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.skip("no GPU")


@pytest.fixture(scope="module")
def base():
    w = config.make_workload(n=2000, r_min=0.2, r_max=0.3, test="uniaxial", seed=21,
                             test_kwargs=dict(v_load_mps=-2.0, dnsave=10, max_steps=5000,
                                              max_strain=0.05))
    mat = autograd.fixed_dt(w["mat"], w["r"])
    return dict(xyz=w["xyz"], r=w["r"], mat=mat, box=w["box"], cfg=w["cfg"], names=NAMES)


def window(base, steps):
    return dict(base, steps=steps)


def theta0(setup):
    m = setup["mat"]
    return np.array([getattr(m, autograd.PARAMS[n][1]) for n in NAMES])


def central(setup, k, rel):
    th = theta0(setup)
    h = rel * abs(th[k])
    tp, tm = th.copy(), th.copy()
    tp[k] += h
    tm[k] -= h
    sp, _ = autograd.run_window(setup, NAMES, tp, sens=False)
    sm, _ = autograd.run_window(setup, NAMES, tm, sens=False)
    return (sp - sm) / (2 * h)


def check_against_fd(setup, rel, tol):
    print(f"case window {setup['steps']} rel {rel}", flush=True)
    _, J = autograd.run_window(setup, NAMES, theta0(setup), sens=True)
    fds = [central(setup, k, rel) for k in range(len(NAMES))]
    scale_E = np.abs(fds[0]).max()
    assert scale_E > 0.0
    for k, name in enumerate(NAMES):
        # friction acts through stick/slip switches: a secant of a few percent
        # crosses them, so its slope is compared on the scale of the main
        # (bond modulus) sensitivity
        floor = 0.1 if name != "friction" else 1.0
        ref = max(np.abs(fds[k]).max(), floor * scale_E)
        err = np.abs(J[:, k, :] - fds[k]).max()
        assert err <= tol * ref, (name, err, ref)


def test_sens_values_equal_the_plain_step(base):
    setup = window(base, 300)
    s_sens, _ = autograd.run_window(setup, NAMES, theta0(setup), sens=True)
    s_plain, _ = autograd.run_window(setup, NAMES, theta0(setup), sens=False)
    assert np.abs(s_plain).max() > 0.0, "platens never loaded the specimen"
    np.testing.assert_array_equal(s_sens, s_plain)


def test_sens_short_window_against_converged_differences(base):
    check_against_fd(window(base, 30), 1e-5, 0.02)


def test_sens_long_window_against_macroscopic_differences(base):
    check_against_fd(window(base, 300), 3e-2, 0.05)


def test_directions_are_independent(base):
    setup = window(base, 300)
    th = theta0(setup)
    _, J4 = autograd.run_window(setup, NAMES, th, sens=True)
    _, J1 = autograd.run_window(setup, NAMES[:1], th[:1], sens=True)
    np.testing.assert_array_equal(J4[:, 0, :], J1[:, 0, :])


def test_autograd_function(base):
    setup = window(base, 300)
    th = torch.tensor(theta0(setup), dtype=torch.float64, requires_grad=True)
    stress = autograd.LoadingResponse.apply(th, setup)
    w = torch.linspace(0.5, 1.5, stress.shape[0], dtype=torch.float64)[:, None]
    loss = (stress * w).sum()
    loss.backward()
    _, J = autograd.run_window(setup, NAMES, theta0(setup), sens=True)
    expect = np.einsum("pe,pke->k", np.broadcast_to(w.numpy(), stress.shape), J * 1e3)
    np.testing.assert_allclose(th.grad.numpy(), expect, rtol=1e-13, atol=0.0)  # J^T g
    fd = (central(setup, 0, 3e-2) * 1e3 * w.numpy()).sum()
    assert abs(th.grad[0].item() - fd) <= 0.02 * abs(fd)


def test_macro_response_against_central_differences(base):
    """Mean axial stress and secant modulus of the 300-step window: the
    torch Function returns macro_from_curve of its run and J^T g; the
    derivatives of both against 3 % central differences, the step the
    long-window test uses, within the bound that test sets on every curve
    point (5 % of the parameter's largest pointwise difference, floored at
    a tenth of the bond modulus'), carried through the mean and through the
    division by the peak strain.  Friction, which acts through stick/slip
    events, is left to that test."""
    setup = window(base, 300)
    th0 = theta0(setup)
    th = torch.tensor(th0, dtype=torch.float64, requires_grad=True)
    out = autograd.MacroResponse.apply(th, setup)
    g = torch.tensor([1.0, 0.5], dtype=torch.float64)
    (out * g).sum().backward()
    stress, J, strain = autograd.run_window(setup, NAMES, th0, sens=True, strain=True)
    v, jac = autograd.macro_from_curve(strain, stress, J)
    scale = np.array([1e3, 1.0])
    np.testing.assert_array_equal(out.detach().numpy(), v * scale)
    np.testing.assert_allclose(th.grad.numpy(), (g.numpy() * scale) @ jac, rtol=1e-13)
    sa = 0.5 * (np.abs(stress[:, 0]) + np.abs(stress[:, 1]))
    eps_peak = -strain[int(np.argmax(sa))]
    fds, pts = [], []
    for k in range(3):
        h = 3e-2 * abs(th0[k])
        tp, tm = th0.copy(), th0.copy()
        tp[k] += h
        tm[k] -= h
        sp, _, ep = autograd.run_window(setup, NAMES, tp, sens=False, strain=True)
        sm, _, em = autograd.run_window(setup, NAMES, tm, sens=False, strain=True)
        fds.append((autograd.macro_from_curve(ep, sp)[0]
                    - autograd.macro_from_curve(em, sm)[0]) / (2 * h))
        pts.append(np.abs(0.5 * (np.abs(sp).sum(axis=1) - np.abs(sm).sum(axis=1)) / (2 * h)).max())
    for k, name in enumerate(NAMES[:3]):
        ref = max(pts[k], 0.1 * pts[0])
        for o, r in ((0, ref), (1, ref / eps_peak)):
            print(f"{name} output {o}: tangent {jac[o, k]:.6e} difference {fds[k][o]:.6e} "
                  f"bound {0.05 * r:.3e}", flush=True)
            assert abs(jac[o, k] - fds[k][o]) <= 0.05 * r, (name, o)


def test_adjoint_run_equals_macro_from_curve(base):
    """rf_adjoint_run (one C-ABI call) gives the values and Jacobian that
    macro_from_curve computes from the same window's curve and tangents."""
    setup = window(base, 300)
    th0 = theta0(setup)
    stress, J, strain = autograd.run_window(setup, NAMES, th0, sens=True, strain=True)
    v, jac = autograd.macro_from_curve(strain, stress, J)
    mat = autograd.with_params(setup["mat"], NAMES, th0)
    c = Context(setup["xyz"], setup["r"], mat, setup["box"])
    c.build_bonds()
    c.set_test(setup["cfg"])
    vals, grad = c.adjoint_run([autograd.PARAMS[n][0] for n in NAMES], setup["steps"])
    np.testing.assert_allclose(vals, v, rtol=1e-12)
    np.testing.assert_allclose(grad, jac, rtol=1e-12, atol=1e-12 * np.abs(jac).max())
    c.close()


def test_derived_dt_refuses_the_parameters_that_move_it():
    """With dt derived from the material (rf_material.dt <= 0), the
    directions that move the time scale setting it are refused (the
    tangents hold dt constant); the others are accepted."""
    from rockfit import abi
    from rockfit._lib import RockfitError
    from rockfit.core import derive

    w = config.make_workload(n=600, r_min=0.2, r_max=0.3, test="uniaxial", seed=3,
                             test_kwargs=dict(v_load_mps=-2.0, dnsave=10))
    assert not w["mat"].dt > 0.0
    d = derive(w["mat"], float(w["r"].min()), float(w["r"].max()))
    moving = (["E_contact", "nu"] if d.T_rayleigh < d.T_bond
              else ["E_bond", "bond_radius_ratio"])
    c = Context(w["xyz"], w["r"], w["mat"], w["box"])
    c.build_bonds()
    c.set_test(w["cfg"])
    for name in moving:
        with pytest.raises(RockfitError) as e:
            c.sens_begin([autograd.PARAMS[name][0]])
        assert e.value.status == abi.RF_ERR_STATE
    c.sens_begin([autograd.PARAMS["friction"][0], autograd.PARAMS["rolling_friction"][0]])
    assert c.sens_step(5) == 5


def test_bond_stretched_past_acut_ends_the_window():
    """A bond stretched past acut, which the plain step drops without a break
    (its prune), ends a sensitivity window with RF_ERR_STATE instead of being
    evaluated on (rf_sens.hip's slot pass)."""
    from rockfit import abi
    from rockfit._lib import RockfitError

    w = config.make_workload(n=600, r_min=0.2, r_max=0.3, test="uniaxial", seed=3)
    mat = autograd.fixed_dt(w["mat"], w["r"])
    # one bonded pair of the sample, on its own
    full = Context(w["xyz"], w["r"], mat, w["box"])
    full.build_bonds()
    st = full.get_state()
    S = full.slots
    q = int(np.nonzero(st["type"] == 1)[0][0])
    pair = [q // S, int(st["nbr"][q])]
    xyz, r = w["xyz"][pair], w["r"][pair]
    xyz, box = config.place_in_box(xyz, r, mat)
    print("case bond stretched past acut", flush=True)

    def stretched():
        c = Context(xyz, r, mat, box)
        c.build_bonds()
        s = c.get_state()
        assert np.count_nonzero(s["type"] == 1) == 2  # one bond, both directions
        R = s["R"].reshape(3, 2)
        d = R[:, 1] - R[:, 0]
        R[:, 1] = R[:, 0] + d / np.linalg.norm(d) * 1.2 * mat.acut
        c.set_state(s)
        return c

    plain = stretched()
    assert plain.step(1) == 1
    assert np.count_nonzero(plain.get_state()["type"] == 1) == 0  # pruned, not broken
    sens = stretched()
    sens.sens_begin([autograd.PARAMS["friction"][0]])
    with pytest.raises(RockfitError) as e:
        sens.sens_step(1)
    assert e.value.status == abi.RF_ERR_STATE

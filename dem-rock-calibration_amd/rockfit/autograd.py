"""The loading response as a differentiable function of material
parameters: torch.autograd.Function over the GPU sensitivity run.

forward(theta) runs one loading window of a fixed sample on the GPU with
the parameter directions seeded (rf_sens_*), returns the platen stresses at
the curve points (MPa, bottom and top) and keeps the Jacobian
d(stress)/d(theta) that the same run produced; backward returns J^T g.
This replaces the finite-difference Jacobian of the reference's calibration
(CalculateStepCG, mp_determineconjgradient.cu:84-142).
"""
import ctypes as C

import numpy as np
import torch

from . import abi
from .core import Context, derive

# name -> (rf_param, Material field); E in reduced units (GPa)
PARAMS = {
    "E_bond": (abi.RF_PARAM_E_BOND, "E_bond"),
    "nu": (abi.RF_PARAM_NU, "nu"),
    "E_contact": (abi.RF_PARAM_E_CONTACT, "E_contact"),
    "friction": (abi.RF_PARAM_FRICTION, "friction"),
    "rolling_friction": (abi.RF_PARAM_ROLLING_FRICTION, "rolling_friction"),
    "bond_radius_ratio": (abi.RF_PARAM_BOND_RADIUS_RATIO, "bond_radius_ratio"),
}


def with_params(mat, names, values):
    """Copy of `mat` with the named fields set; dt fixed to the base value
    (the step size is a constant of a sensitivity run)."""
    m = abi.Material.from_buffer_copy(mat)
    if not (m.dt > 0.0):
        raise ValueError("fix dt first (fixed_dt): the step size must not follow the parameters")
    for nme, v in zip(names, values):
        setattr(m, PARAMS[nme][1], float(v))
    return m


def fixed_dt(mat, r):
    """Material with dt frozen at the value rf_derive gives for it."""
    m = abi.Material.from_buffer_copy(mat)
    if not (m.dt > 0.0):
        m.dt = derive(mat, float(np.min(r)), float(np.max(r))).dt
    return m


def run_window(setup, names, values, sens=True, strain=False):
    """One loading window: returns (stress (points, 2) reduced units,
    Jacobian (points, nparam, 2) or None), and the curve's strain column
    (points,) as a third value when `strain` is set."""
    mat = with_params(setup["mat"], names, values)
    c = Context(setup["xyz"], setup["r"], mat, setup["box"])
    c.build_bonds()
    c.set_test(setup["cfg"])
    if sens:
        c.sens_begin([PARAMS[n][0] for n in names])
        done = c.sens_step(setup["steps"])
    else:
        done = c.step(setup["steps"])
    if done != setup["steps"]:
        raise RuntimeError(f"the loading window ended after {done} of {setup['steps']} steps")
    curve = c.curve()
    J = c.sens_curve() if sens else None
    if strain:
        return curve[:, 1:3].copy(), J, curve[:, 0].copy()
    return curve[:, 1:3].copy(), J


def run_frozen(setup, names, values, recorded=None):
    """One loading window of the frozen-branch model (rf_sens_branches).

    recorded None: the window decides and records every discrete decision
    (bond breaks, contact touch and slip, rolling and its chatter, platen
    contacts, the platen velocity); else it replays the decisions the
    context `recorded` recorded.  Returns (stress (points, 2), Jacobian
    (points, nparam, 2), strain (points,), context); keep the recording
    context alive while windows replay it."""
    mat = with_params(setup["mat"], names, values)
    c = Context(setup["xyz"], setup["r"], mat, setup["box"])
    c.build_bonds()
    c.set_test(setup["cfg"])
    if recorded is None:
        c.sens_branches(abi.RF_BRANCH_RECORD, setup["steps"])
    else:
        c.sens_branches(abi.RF_BRANCH_REPLAY, 0, recorded)
    c.sens_begin([PARAMS[n][0] for n in names])
    done = c.sens_step(setup["steps"])
    if done != setup["steps"]:
        raise RuntimeError(f"the loading window ended after {done} of {setup['steps']} steps")
    curve = c.curve()
    return curve[:, 1:3].copy(), c.sens_curve(), curve[:, 0].copy(), c


def macro_from_curve(strain, stress, J=None, peak=None):
    """Mean axial stress and secant modulus of a loading window (reduced
    units: stress and modulus in GPa), and their Jacobian when the curve's
    is given.

    strain (points,) is the curve's platen strain, (size - size0) / size0,
    negative in compression; stress (points, 2) holds the bottom and top
    platen stresses; J (points, nparam, 2) their derivatives.  The axial
    stress of a point is (|s_bottom| + |s_top|) / 2 (SumForcesULoading,
    md_results.cu:56-99); the mean runs over the window's points.  The
    secant modulus is s_peak / e_peak (md_results.cu:149), the peak being the
    first point of highest axial stress (the peak moves only on a strict
    increase, md_results.cu:70-75) and e_peak = -strain there.  The peak's
    position is an event: its derivative follows the point it picks; `peak`
    holds it at a given point (the frozen-branch model).
    Returns (values (2,), Jacobian (2, nparam) or None)."""
    sa = 0.5 * (np.abs(stress[:, 0]) + np.abs(stress[:, 1]))
    p = int(np.argmax(sa)) if peak is None else int(peak)
    eps = -float(strain[p])
    if not eps > 0.0:
        raise ValueError("the window's highest stress comes before any platen travel")
    vals = np.array([sa.mean(), sa[p] / eps])
    if J is None:
        return vals, None
    dsa = 0.5 * (np.sign(stress[:, 0])[:, None] * J[:, :, 0]
                 + np.sign(stress[:, 1])[:, None] * J[:, :, 1])
    return vals, np.stack([dsa.mean(axis=0), dsa[p] / eps])


class LoadingResponse(torch.autograd.Function):
    """stress[MPa] (points, 2) = f(theta); setup = dict(xyz, r, mat (dt > 0),
    box, cfg, steps, names)."""

    @staticmethod
    def forward(ctx, theta, setup):
        values = theta.detach().cpu().numpy().astype(np.float64)
        stress, J = run_window(setup, setup["names"], values, sens=True)
        if J.shape[0] != stress.shape[0]:
            raise RuntimeError("curve and curve tangents differ in length")
        ctx.save_for_backward(torch.from_numpy(J * 1e3))
        return torch.from_numpy(stress * 1e3).to(theta.dtype)

    @staticmethod
    def backward(ctx, g):
        (J,) = ctx.saved_tensors
        grad = torch.einsum("pe,pke->k", g.to(J.dtype), J)
        return grad, None


class MacroResponse(torch.autograd.Function):
    """[mean axial stress (MPa), secant modulus (GPa)] of a loading window
    as a function of theta (macro_from_curve over the curve of one GPU
    sensitivity run); setup as for LoadingResponse."""

    @staticmethod
    def forward(ctx, theta, setup):
        values = theta.detach().cpu().numpy().astype(np.float64)
        stress, J, strain = run_window(setup, setup["names"], values, sens=True, strain=True)
        v, jac = macro_from_curve(strain, stress, J)
        scale = np.array([1e3, 1.0])
        ctx.save_for_backward(torch.from_numpy(jac * scale[:, None]))
        return torch.from_numpy(v * scale).to(theta.dtype)

    @staticmethod
    def backward(ctx, g):
        (jac,) = ctx.saved_tensors
        return g.to(jac.dtype) @ jac, None

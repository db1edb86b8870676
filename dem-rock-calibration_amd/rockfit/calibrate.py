"""Calibration on top of the GPU loading step.

Two drivers sit above the C-ABI:

- the reference's own: the conjugate-gradient step over finite differences
  (DetermineConjGradient, mp_determineconjgradient.cu:146-368, mp_task.h:233-311)
  is `RockFit-DEM DCG` (cli/rf_cg.h), whose difference runs go to the GPU
  as one batch; the surrogate farm (NN_MP.py) drives `RockFit-DEM M(...)`
  (rockfit.farm) and scores results with `target_functional` below;
- gradient ones: `ConjugateGradient` is the same CG step
  (CalculateNewParameters2, mp_task.h:233-311) with the response's Jacobian
  given, and `cg_step_gradient` feeds it the forward-mode tangents of one GPU
  run (rf_sens_*) instead of NParam + 1 difference runs; `fit_curve`
  matches a loading curve with Gauss-Newton / Levenberg-Marquardt steps on
  the same tangents.
"""
import numpy as np

from . import autograd

# ResultCalculateOneSample fields compared by the farm's objective
# (NN_MP.py:849): strain_u, stress_u, linearity_u, destruction_u, stress_b,
# destruction_b.
FARM_FEATURES = (6, 7, 8, 9, 11, 13)


def farm_features(row):
    """The six compared quantities of one ResultCalculateOneSample row (14
    values)."""
    row = np.asarray(row, dtype=np.float64)
    return row[list(FARM_FEATURES)]


def target_functional(points, target, weights):
    """ReCalculateTargetFunctional (NN_MP.py:461-504): per point the sum over
    six quantities of weight * d^2 with
      d = x/t - 1 if x >= t else t/x - 1       (quantities 0, 1, 2, 4),
      d = 10 if 1 - x < t or 1 - x > 0.7 else 0 (quantity 3, destruction),
      d = 10 if 1 - x < t else 0               (quantity 5),
    each d capped at 10."""
    out = []
    for x in points:
        y = 0.0
        for j in range(6):
            if j == 3:
                d = 10.0 if (1.0 - x[j] < target[j] or 1.0 - x[j] > 0.7) else 0.0
            elif j == 5:
                d = 10.0 if 1.0 - x[j] < target[j] else 0.0
            elif x[j] >= target[j]:
                d = x[j] / target[j] - 1.0
            else:
                d = target[j] / x[j] - 1.0
            if d > 10:
                d = 10.0
            y += weights[j] * d * d
        out.append(y)
    return out


def cg_functional(results, goal, weight_mult):
    """The CG's weighted squares W_j (G_j - R_j)^2 with W_j = mult_j / G_j^2
    (CalculateStartParameters and CalculateNewParameters2, mp_task.h:138-139,
    275-279)."""
    goal = np.asarray(goal, dtype=np.float64)
    W = np.asarray(weight_mult, dtype=np.float64) / goal ** 2
    D = goal - np.asarray(results, dtype=np.float64)
    return W * D * D


def fit_curve(setup, target, names, theta0, iters=10, lam=1e-3, tol=1e-10, scale=None,
              verbose=False):
    """Fit material parameters to a loading curve on the GPU.

    setup: dict(xyz, r, mat (dt fixed), box, cfg, steps) as for
    autograd.run_window; target: (points, 2) platen stresses in reduced
    units; names: parameters (autograd.PARAMS) started at theta0.
    Levenberg-Marquardt in the scaled variables theta / scale: each
    iteration is one GPU run that returns the curve and its tangents.
    Returns (theta, history) with history = [(loss, theta), ...]."""
    theta = np.asarray(theta0, dtype=np.float64).copy()
    scale = np.abs(theta) if scale is None else np.asarray(scale, dtype=np.float64)
    target = np.asarray(target, dtype=np.float64)
    stress, J = autograd.run_window(setup, names, theta, sens=True)
    r = (stress - target).ravel()
    loss = float(r @ r)
    history = [(loss, theta.copy())]
    for it in range(iters):
        # J: (points, nparam, 2) -> residual rows (points*2, nparam), scaled
        A = np.transpose(J, (0, 2, 1)).reshape(-1, len(names)) * scale
        H = A.T @ A
        g = A.T @ r
        while True:
            step = np.linalg.solve(H + lam * np.diag(np.diag(H)), -g)
            trial = theta + step * scale
            s2, J2 = autograd.run_window(setup, names, trial, sens=True)
            r2 = (s2 - target).ravel()
            loss2 = float(r2 @ r2)
            if verbose:
                print(f"iter {it} lam {lam:.1e} loss {loss:.6e} -> {loss2:.6e} theta {trial}")
            if loss2 < loss:
                theta, stress, J, r, loss = trial, s2, J2, r2, loss2
                lam = max(lam / 10.0, 1e-12)
                break
            lam *= 10.0
            if lam > 1e8:
                return theta, history
        history.append((loss, theta.copy()))
        if loss <= tol * history[0][0]:
            break
    return theta, history


class ConjugateGradient:
    """The reference's conjugate-gradient calibration step for any number of
    parameters and results, with the response's Jacobian given.

    CalculateStartParameters (mp_task.h:134-151): weights W_j = mult_j /
    goal_j^2.  CalculateNewParameters2 (mp_task.h:233-311), normalisation
    coefficients 1: D = goal - R, F_j = W_j D_j^2, gradient
    g_i = sum_j -2 W_j D_j dR_ij, direction d = g + beta g_old with
    beta = |g|^2 / |g_old|^2 (the reference adds beta times the old
    gradient, its direction array having been overwritten with it), step
    t = -a1 / a0 with a1 = sum_j W_j D_j (dR d)_j and
    a0 = sum_j W_j (dR d)_j^2 (the Gauss-Newton minimum along d), and
    params - t d clipped at `lower` (SetMaterialContractions).  The same
    arithmetic, in the same loop order, as cli/rf_cg.h's Task::update, whose
    dR comes from forward differences (mp_determineconjgradient.cu:84-142).
    dR[i][j] = d R_j / d param_i."""

    def __init__(self, goal, weight_mult, lower=None):
        self.goal = np.asarray(goal, dtype=np.float64)
        self.weight = np.asarray(weight_mult, dtype=np.float64) / (self.goal * self.goal)
        self.lower = None if lower is None else np.asarray(lower, dtype=np.float64)
        self.dF = None
        self.dd = None
        self.F = None
        self.t = 0.0

    def step(self, params, R, dR):
        params = np.asarray(params, dtype=np.float64)
        R = np.asarray(R, dtype=np.float64)
        dR = np.asarray(dR, dtype=np.float64)
        npar, nres = dR.shape
        if self.dF is None:
            self.dF = np.zeros(npar)
            self.dd = np.zeros(npar)
        D = [self.goal[j] - R[j] for j in range(nres)]
        self.F = np.array([self.weight[j] * D[j] * D[j] for j in range(nres)])
        g_old = 0.0
        g_new = 0.0
        dd = np.zeros(npar)
        dF = np.zeros(npar)
        for i in range(npar):
            dd[i] = self.dF[i]
            g_old += self.dF[i] * self.dF[i]
            g = 0.0
            for j in range(nres):
                g += -2 * self.weight[j] * D[j] * dR[i][j]
            dF[i] = g
            g_new += g * g
        beta = g_new / g_old if g_old > 1e-12 else 0.0
        for i in range(npar):
            dd[i] = dF[i] + beta * dd[i]
        a0 = 0.0
        a1 = 0.0
        for j in range(nres):
            for i in range(npar):
                a1 += self.weight[j] * D[j] * dR[i][j] * dd[i]
                for k in range(npar):
                    a0 += self.weight[j] * dR[i][j] * dd[i] * dR[k][j] * dd[k]
        self.t = -a1 / a0
        self.dF, self.dd = dF, dd
        out = np.array([params[i] - self.t * dd[i] for i in range(npar)])
        if self.lower is not None:
            out = np.maximum(out, self.lower)
        return out


def window_response(setup, names, theta, frozen=None):
    """Mean axial stress and secant modulus of a loading window (reduced
    units) and their Jacobian dR[i][j] = d R_j / d theta_i, from one GPU
    run carrying the tangents.  frozen: a context that recorded the
    window's decisions (autograd.run_frozen); the run then replays them and
    the secant is taken at the recorded peak point."""
    if frozen is None:
        stress, J, strain = autograd.run_window(setup, names, theta, sens=True, strain=True)
        v, jac = autograd.macro_from_curve(strain, stress, J)
        return v, jac.T
    ctx, peak = frozen
    stress, J, strain, c = autograd.run_frozen(setup, names, theta, recorded=ctx)
    c.close()
    v, jac = autograd.macro_from_curve(strain, stress, J, peak=peak)
    return v, jac.T


def cg_step_gradient(cg, setup, names, theta, frozen=None):
    """One CG step whose Jacobian is the tangents of a single GPU run
    (NParam directions carried together): returns (new theta, R)."""
    R, dR = window_response(setup, names, theta, frozen)
    return cg.step(theta, R, dR), R


def cg_step_differences(cg, setup, names, theta, rel, frozen=None):
    """The same CG step with the reference's Jacobian: forward differences,
    parameter i moved by rel * theta_i, NParam + 1 runs
    (mp_determineconjgradient.cu:84-142).  Returns (new theta, R)."""
    theta = np.asarray(theta, dtype=np.float64)
    R0, _ = window_response(setup, names, theta, frozen)
    dR = np.zeros((len(theta), len(R0)))
    for i in range(len(theta)):
        tp = theta.copy()
        h = rel * theta[i]
        tp[i] += h
        Ri, _ = window_response(setup, names, tp, frozen)
        dR[i] = (Ri - R0) / h
    return cg.step(theta, R0, dR), R0

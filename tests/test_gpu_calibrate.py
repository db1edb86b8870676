"""Gradient calibration on the GPU (rockfit.calibrate.fit_curve): a loading
curve made with known parameters is matched again from a start 20-25 % off,
with Levenberg-Marquardt steps whose Jacobian is the forward-mode tangent of
the same run.  Contact switches (stick/slip, rolling) make the response
rough at the 1e-6 level of the initial misfit, which bounds how closely the
parameters can be recovered (here 1e-4 .. 1e-3 relative).  Sample and window as tests/test_gpu_sens.py (2000 spheres,
platens load the specimen to about 3 MPa in 300 steps)."""
import numpy as np
import pytest

from conftest import gpu_available
from rockfit import autograd, calibrate, config

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.skip("no GPU")


@pytest.fixture(scope="module")
def setup():
    w = config.make_workload(n=2000, r_min=0.2, r_max=0.3, test="uniaxial", seed=21,
                             test_kwargs=dict(v_load_mps=-2.0, dnsave=10, max_steps=5000,
                                              max_strain=0.05))
    mat = autograd.fixed_dt(w["mat"], w["r"])
    return dict(xyz=w["xyz"], r=w["r"], mat=mat, box=w["box"], cfg=w["cfg"], steps=300)


@pytest.mark.parametrize("names,factors,tols", [
    (["E_bond"], [0.8], [1e-3]),
    (["E_bond", "nu"], [0.8, 1.25], [1e-3, 2e-2]),
])
def test_fit_recovers_parameters(setup, names, factors, tols):
    print(f"case fit {names}", flush=True)
    m = setup["mat"]
    truth = np.array([getattr(m, autograd.PARAMS[n][1]) for n in names])
    target, _ = autograd.run_window(setup, names, truth, sens=False)
    assert np.abs(target).max() > 0.0
    theta0 = truth * np.array(factors)
    theta, hist = calibrate.fit_curve(setup, target, names, theta0, iters=12)
    losses = [h[0] for h in hist]
    print("losses", losses, "theta", theta, "truth", truth, flush=True)
    assert all(b < a for a, b in zip(losses, losses[1:]))
    # contact switches make the response rough below ~1e-6 of the initial
    # misfit: the first steps converge quadratically, then the loss levels off
    assert losses[-1] < 1e-5 * losses[0]
    for k in range(len(names)):
        assert abs(theta[k] - truth[k]) <= tols[k] * abs(truth[k]), (names[k], theta[k], truth[k])

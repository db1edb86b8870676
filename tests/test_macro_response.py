"""The reduction of a loading window's curve to mean axial stress and secant
modulus (rockfit.autograd.macro_from_curve), and its Jacobian, on synthetic
curves with closed forms.  CPU only: the reduction is host arithmetic over
the curve and curve tangents a GPU sensitivity run returns."""
import numpy as np
import pytest

from rockfit.autograd import macro_from_curve

P = 60
EPS = -np.arange(P) * 1e-5  # curve strain column: negative in compression


def curve(theta, imbalance=0.98):
    """Platen stresses s_top = a|e| + b e^2, s_bottom = -(imbalance a|e| + b e^2)
    and their exact derivatives with respect to (a, b)."""
    a, b = theta
    e = np.abs(EPS)
    top = a * e + b * e * e
    bot = -(imbalance * a * e + b * e * e)
    J = np.zeros((P, 2, 2))
    J[:, 0, 0], J[:, 1, 0] = -imbalance * e, -e * e
    J[:, 0, 1], J[:, 1, 1] = e, e * e
    return np.stack([bot, top], axis=1), J


def test_linear_curve_closed_form():
    a = 30.0
    stress, J = curve((a, 0.0))
    v, jac = macro_from_curve(EPS, stress, J)
    e = np.abs(EPS)
    assert v[0] == pytest.approx(0.99 * a * e.mean(), rel=1e-14)
    assert v[1] == pytest.approx(0.99 * a, rel=1e-14)  # peak at the last point
    assert jac[0, 0] == pytest.approx(0.99 * e.mean(), rel=1e-14)
    assert jac[1, 0] == pytest.approx(0.99, rel=1e-14)
    assert jac[1, 1] == pytest.approx(e[-1], rel=1e-14)  # d/db of (a e + b e^2)/e


def test_jacobian_against_central_differences_with_an_interior_peak():
    theta = np.array([30.0, -4e4])  # peak near e = a / (2|b|) = 3.75e-4, point 37
    stress, J = curve(theta)
    v, jac = macro_from_curve(EPS, stress, J)
    sa = 0.5 * (np.abs(stress[:, 0]) + np.abs(stress[:, 1]))
    p = int(np.argmax(sa))
    assert 0 < p < P - 1
    assert v[1] == pytest.approx(sa[p] / abs(EPS[p]), rel=1e-15)
    for k in range(2):
        h = 1e-6 * abs(theta[k])
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        vp, _ = macro_from_curve(EPS, curve(tp)[0])
        vm, _ = macro_from_curve(EPS, curve(tm)[0])
        np.testing.assert_allclose(jac[:, k], (vp - vm) / (2 * h), rtol=1e-7)


def test_first_of_equal_peaks_is_the_peak():
    sa = np.array([0.0, 1.0, 2.0, 2.0, 1.0])
    strain = -np.array([0.0, 1.0, 2.0, 3.0, 4.0]) * 1e-3
    stress = np.stack([-sa, sa], axis=1)
    v, _ = macro_from_curve(strain, stress)
    assert v[1] == 2.0 / 2e-3
    assert v[0] == pytest.approx(sa.mean(), rel=1e-15)


def test_unloaded_window_is_an_error():
    with pytest.raises(ValueError):
        macro_from_curve(EPS, np.zeros((P, 2)))

"""Curve post-processing of the C-ABI (rf_linearity_deviation,
rf_curve_moduli) on synthetic curves whose answers are closed forms.

Curve layout as rf_get_curve: strain (negative in compression), bottom
stress (negative), top stress; the stress is s = (top - bottom) / 2."""
import numpy as np
import pytest

from rockfit import core


def curve_of(eps, sig):
    return np.stack([-eps, -sig, sig], axis=1)


def test_linear_curve():
    E = 17.0
    eps = np.linspace(0.0, 0.01, 501)
    sig = E * eps
    c = curve_of(eps, sig)
    assert core.linearity_deviation(c, eps[-1], sig[-1]) == pytest.approx(1.0, abs=1e-12)
    m = core.curve_moduli(c, eps[-1], sig[-1], 0.1, 0.5)
    assert m.secant == pytest.approx(E, rel=1e-14)
    assert m.lsq == pytest.approx(E, rel=1e-12)
    assert abs(m.lsq_intercept) < 1e-12
    assert m.r2 == pytest.approx(1.0, abs=1e-12)
    assert m.points == np.count_nonzero((sig >= 0.1 * sig[-1]) & (sig <= 0.5 * sig[-1]))


def test_parabola_linearity_deviation_closed_form():
    """s = E eps - c eps^2 peaks at eps* = E/2c, s* = c eps*^2; its secant is
    E/2, and the deviation integral int_0^eps* (A eps - s)^2 deps
    / (eps* s*^2) = 1/30, so LD -> 1 + 1/30 as the sampling refines."""
    E, c = 20.0, 1000.0
    ep = E / (2 * c)
    sp = c * ep * ep
    for n, tol in ((400, 2e-2), (4000, 2e-3)):
        eps = np.linspace(0.0, 1.3 * ep, n)
        sig = E * eps - c * eps * eps
        ld = core.linearity_deviation(curve_of(eps, sig), ep, sp)
        assert ld - 1.0 == pytest.approx(1.0 / 30.0, rel=tol)


def test_least_squares_modulus_on_ascending_branch():
    """Only samples before the peak and inside the stress window enter the
    fit; it equals numpy's least-squares line through them."""
    E, c = 20.0, 1000.0
    ep = E / (2 * c)
    eps = np.linspace(0.0, 1.8 * ep, 3001)
    sig = E * eps - c * eps * eps
    sp = sig.max()
    m = core.curve_moduli(curve_of(eps, sig), ep, sp, 0.2, 0.6)
    ipk = int(np.argmax(sig >= sp))
    sel = np.arange(len(eps)) <= ipk
    sel &= (sig >= 0.2 * sp) & (sig <= 0.6 * sp)
    slope, icpt = np.polyfit(eps[sel], sig[sel], 1)
    assert m.points == sel.sum()
    assert m.lsq == pytest.approx(slope, rel=1e-10)
    assert m.lsq_intercept == pytest.approx(icpt, rel=1e-8, abs=1e-14)
    assert m.secant == pytest.approx(sp / ep, rel=1e-14)
    assert 0.98 < m.r2 < 1.0


def test_samples_past_n_read_as_zero():
    """The loop reads one sample past the last stored one, which counts as
    zero (the reference's zero-filled buffer): appending explicit zero rows
    changes nothing."""
    eps = np.linspace(0.0, 0.004, 50)
    sig = 20.0 * eps - 1000.0 * eps ** 2
    c = curve_of(eps, sig)
    a = core.linearity_deviation(c, 0.01, 0.1)
    b = core.linearity_deviation(np.vstack([c, np.zeros((3, 3))]), 0.01, 0.1)
    # with the extra rows the loop may run further, but only over zero terms
    assert a == b


def test_bad_arguments():
    from rockfit._lib import RockfitError
    with pytest.raises(RockfitError):
        core.linearity_deviation(np.zeros((3, 3)), 0.0, 1.0)
    with pytest.raises(RockfitError):
        core.curve_moduli(np.zeros((3, 3)), 0.01, 1.0, 0.6, 0.2)

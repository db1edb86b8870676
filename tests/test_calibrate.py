"""Host parts of the calibration drivers (rockfit.calibrate): the farm's
objective (NN_MP.py:461-504, 849) and the CG functional
(mp_task.h:138-139, 275-279), on hand-computed cases.  The reference holds
no fixtures for them; each case below is worked out from the rules the
docstrings state."""
import numpy as np
import pytest

from rockfit import calibrate

TARGET = [0.005, 100e6, 1.0, 0.1, 10e6, 0.05]
WEIGHTS = [1, 1, 5000, 1, 1, 1]


def test_objective_relative_errors():
    # x >= t: x/t - 1; x < t: t/x - 1; destruction terms inside their bounds
    x = [0.01, 50e6, 1.2, 0.5, 20e6, 0.5]
    (y,) = calibrate.target_functional([x], TARGET, WEIGHTS)
    assert y == pytest.approx(1.0 + 1.0 + 5000 * 0.2 ** 2 + 0.0 + 1.0 + 0.0, rel=1e-12)


def test_objective_penalties_and_cap():
    # strain 200x the target is capped at 10; destruction left 0.05 < 0.1 and
    # Brazilian destruction left 0.03 < 0.05 are penalised with 10 each
    x = [1.0, 100e6, 1.0, 0.95, 10e6, 0.97]
    (y,) = calibrate.target_functional([x], TARGET, WEIGHTS)
    assert y == pytest.approx(100.0 + 0.0 + 0.0 + 100.0 + 0.0 + 100.0, rel=1e-12)
    # too much destruction (1 - 0.2 > 0.7) is penalised as well
    x = [0.005, 100e6, 1.0, 0.2, 10e6, 0.5]
    (y,) = calibrate.target_functional([x], TARGET, WEIGHTS)
    assert y == pytest.approx(100.0, rel=1e-12)
    # on target: zero
    x = [0.005, 100e6, 1.0, 0.5, 10e6, 0.5]
    assert calibrate.target_functional([x], TARGET, WEIGHTS) == [0.0]


def test_farm_features():
    row = np.arange(14, dtype=float)
    np.testing.assert_array_equal(calibrate.farm_features(row), [6, 7, 8, 9, 11, 13])


def test_cg_functional():
    F = calibrate.cg_functional([0.004, 0.1, 1.0, 1.1], [0.005, 0.104, 1.0, 1.0],
                                [50, 100, 100, 10000])
    np.testing.assert_allclose(F, [50 * 0.2 ** 2, 100 * (0.004 / 0.104) ** 2, 0.0,
                                   10000 * 0.1 ** 2], rtol=1e-12)

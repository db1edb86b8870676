"""The post-processing rf_run reports (linearity deviation, destruction)
equals the host functions applied to the curve and bond counts the same run
returns; the moduli of the run's curve are well defined (this fast,
impact-like loading gives a noisy ascending branch, so no closed form is
asserted here: the synthetic curves of tests/test_post.py pin the fit)."""
import numpy as np
import pytest

from conftest import gpu_available
from rockfit import config, core
from rockfit._lib import RockfitError
from rockfit.core import Context

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.skip("no GPU")


def test_run_post_processing_equals_host_functions():
    print("case post n=2000 weak bonds to fracture", flush=True)
    w = config.make_workload(n=2000, r_min=0.2, r_max=0.3, test="uniaxial", seed=4,
                             mat_kwargs=dict(sigma_mpa=8.0, tau_mpa=12.0),
                             test_kwargs=dict(v_load_mps=-2.0, dnsave=5, max_steps=20000,
                                              max_strain=0.05))
    g = Context(w["xyz"], w["r"], w["mat"], w["box"])
    g.build_bonds()
    g.set_test(w["cfg"])
    ts = g.run()
    assert ts.done and ts.fracture == 1
    curve = g.curve()
    assert len(curve) == ts.n_curve
    ld = core.linearity_deviation(curve, ts.peak_strain, ts.peak_stress)
    assert ld == ts.linearity_deviation
    bc = g.bond_counts()
    assert ts.destruction == bc[ts.n_curve - 1] / bc[0]
    m = core.curve_moduli(curve, ts.peak_strain, ts.peak_stress, 0.1, 0.5)
    assert m.points >= 3
    assert m.secant == ts.peak_stress / ts.peak_strain
    assert np.isfinite(m.lsq) and 0.0 <= m.r2 <= 1.0
    g.close()


@pytest.mark.parametrize("kind", ["uniaxial", "brazilian"])
def test_run_entry_points_equal_set_test_and_run(kind):
    """rf_run_uniaxial / rf_run_brazilian give the results of rf_set_test +
    rf_run on the same built sample, bit for bit, and refuse a platen
    configuration of the other test."""
    print(f"case run entry point {kind} n=1500", flush=True)
    shape, aspect = ("cylinder", 2.0) if kind == "uniaxial" else ("disc", 0.5)
    w = config.make_workload(n=1500, r_min=0.2, r_max=0.3, test=kind, shape=shape, aspect=aspect,
                             seed=6, mat_kwargs=dict(sigma_mpa=8.0, tau_mpa=12.0),
                             test_kwargs=dict(v_load_mps=-2.0, dnsave=5, max_steps=8000))
    a = Context(w["xyz"], w["r"], w["mat"], w["box"])
    a.build_bonds()
    res = a.run_uniaxial(w["cfg"]) if kind == "uniaxial" else a.run_brazilian(w["cfg"])
    b = Context(w["xyz"], w["r"], w["mat"], w["box"])
    b.build_bonds()
    b.set_test(w["cfg"])
    ts = b.run()
    assert res.peak_stress == ts.peak_stress and res.peak_strain == ts.peak_strain
    assert res.secant_modulus == ts.peak_stress / ts.peak_strain
    assert res.linearity_deviation == ts.linearity_deviation
    assert res.destruction == ts.destruction
    assert (res.steps, res.curve_points, res.fracture, res.kind) == (ts.step, ts.n_curve,
                                                                     ts.fracture, ts.kind)
    other = a.run_brazilian if kind == "uniaxial" else a.run_uniaxial
    with pytest.raises(RockfitError):
        other(w["cfg"])
    a.close()
    b.close()

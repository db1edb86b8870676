"""Batched samples (rf_batch_*) against the same samples run alone.

Pin: a sample advanced inside a batch ends bit for bit where it ends when it
runs alone (and therefore where the oracle ends).  Samples differ in size,
material, test kind and skin, so they halt for list rebuilds at different
steps and finish at different steps.
"""
import numpy as np
import pytest

import pyoracle
from conftest import gpu_available
from helpers import compare_states, compare_test_states
from rockfit import config
from rockfit.core import Batch, Context

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.skip("no GPU")


def workloads():
    return [
        config.make_workload(n=900, r_min=0.2, r_max=0.3, test="uniaxial", seed=11,
                             test_kwargs=dict(v_load_mps=-2.0, dnsave=5, max_steps=700,
                                              max_strain=0.05)),
        config.make_workload(n=1700, r_min=0.2, r_max=0.3, test="uniaxial", seed=12,
                             mat_kwargs=dict(sigma_mpa=8.0, tau_mpa=12.0),
                             test_kwargs=dict(v_load_mps=-2.0, dnsave=1, max_steps=450,
                                              max_strain=0.05)),
        config.make_workload(n=1300, r_min=0.2, r_max=0.3, shape="disc", aspect=0.5,
                             test="brazilian", seed=13,
                             test_kwargs=dict(v_load_mps=-2.0, dnsave=5, max_steps=600)),
        config.make_workload(n=600, r_min=0.2, r_max=0.35, test="uniaxial", seed=14,
                             mat_kwargs=dict(E_bond_gpa=35.0),
                             test_kwargs=dict(v_load_mps=-3.0, dnsave=3, max_steps=500,
                                              max_strain=0.05)),
    ]


def contexts(ws, skins):
    out = []
    for w, sk in zip(ws, skins):
        c = Context(w["xyz"], w["r"], w["mat"], w["box"], skin=sk)
        c.build_bonds()
        c.set_test(w["cfg"])
        out.append(c)
    return out


@pytest.mark.parametrize("form", [-1, 1, 3, 4])
def test_batch_steps_equal_single_runs(form):
    """form -1: the autotuner (every form runs during its first chunks);
    1: one shared launch, blocks interleaved; 3: samples one after another;
    4: samples on concurrent streams."""
    ws = workloads()
    skins = [-1.0, 0.02, -1.0, 0.01]
    batch_ctx = contexts(ws, skins)
    single_ctx = contexts(ws, skins)
    b = Batch(batch_ctx)
    if form >= 0:
        b.set_form(form)
        assert b.tuning()[0] == form
    total = [0] * len(ws)
    for chunk in (1, 37, 300):
        print(f"batch form {form} chunk {chunk}", flush=True)
        done = b.step(chunk)
        for k, c in enumerate(single_ctx):
            assert c.step(chunk) == done[k]
            total[k] += done[k]
    for k, (cb, cs) in enumerate(zip(batch_ctx, single_ctx)):
        compare_states(cb.get_state(), cs.get_state(), cb.n, cb.slots, what=f"sample {k}")
        compare_test_states(cb.test_state(), cs.test_state(), what=f"sample {k}")
        np.testing.assert_array_equal(cb.curve(), cs.curve())
        np.testing.assert_array_equal(cb.bond_counts(), cs.bond_counts())
    assert any(c.counters().rebuilds > 1 for c in batch_ctx)
    # one sample also against the oracle
    w = ws[1]
    o = pyoracle.Oracle(w["xyz"], w["r"], w["mat"], w["box"], threads=4)
    o.build_bonds()
    o.set_test(w["cfg"])
    assert o.step(total[1]) == total[1]
    compare_states(batch_ctx[1].get_state(), o.get_state(), o.n, 32, what="batch vs oracle")


def test_batch_run_equals_single_runs():
    ws = workloads()
    skins = [-1.0] * len(ws)
    b = Batch(contexts(ws, skins))
    res = b.run()
    assert b.last_ms() > 0.0
    order, ms = b.tuning()
    print(f"block order {order}, ms per step {ms}", flush=True)
    assert order in (-1, 0, 1, 2, 3, 4)
    if order >= 0:
        assert all(m > 0.0 for m in ms) and ms[order] == min(ms)
    for k, c in enumerate(contexts(ws, skins)):
        t = c.run()
        for f in ("peak_stress", "peak_strain", "linearity_deviation", "destruction", "step"):
            assert getattr(res[k], f) == getattr(t, f), (k, f, getattr(res[k], f), getattr(t, f))
        assert res[k].done == 1

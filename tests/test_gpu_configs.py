"""The loading configurations through the HIP path, against the CPU oracle.

C1: 5k-sphere cylinder, strain rate 0.1/s, 2000 loading steps.
C2: the 100k-sphere benchmark sample, platens at 2 m/s for 2000 steps
    (bonds break, pairs touch).
C1 whole test: the same sample and material loaded at 1 m/s through
    rf_run_uniaxial to its end, macroscopic results against the oracle.
C3: 50k-sphere Brazilian disc, platens at 0.02 m/s, 500 steps.
C4: 64 samples (N log-uniform in [2k, 40k], Latin-hypercube material) as one
    batch for 200 steps: every sample equal to its own single run, two of
    them equal to the oracle.
C5: the 20k-sphere elastic window of the gradients (2000 steps to 0.1 %
    strain, dt fixed), forward run.
All samples come from the random-sequential-insertion + FIRE generator.
Bar: lists, bond numbering and breaks identical; state, histories,
controller, curve and bond counts bitwise equal.
"""
import numpy as np
import pytest

import pyoracle
from helpers import compare_states, compare_test_states
from rockfit import config
from rockfit.core import Batch, Context

pytestmark = pytest.mark.gpu

THREADS = 16


def _run_pair(w, chunks, what):
    g = Context(w["xyz"], w["r"], w["mat"], w["box"])
    o = pyoracle.Oracle(w["xyz"], w["r"], w["mat"], w["box"], threads=THREADS)
    g.build_bonds()
    o.build_bonds()
    compare_states(g.get_state(), o.get_state(), g.n, g.slots, what=f"{what} build")
    g.set_test(w["cfg"])
    o.set_test(w["cfg"])
    done = 0
    for chunk in chunks:
        kg, ko = g.step(chunk), o.step(chunk)
        assert kg == ko == chunk, (kg, ko)
        done += chunk
        print(f"{what}: {done} steps equal", flush=True)
        compare_states(g.get_state(), o.get_state(), g.n, g.slots, what=f"{what} step {done}")
        compare_test_states(g.test_state(), o.test_state(), what=f"{what} step {done}")
    np.testing.assert_array_equal(g.curve(), o.curve())
    np.testing.assert_array_equal(g.bond_counts(), o.bond_counts())
    return g, o


@pytest.mark.timeout(300)
def test_c1_uniaxial_5k_strain_rate():
    print("case C1 5k uniaxial 0.1/s", flush=True)
    w = config.config_workload("C1")
    g, _ = _run_pair(w, (1, 499, 1500), "C1")
    ts = g.test_state()
    assert ts.step == 2000 and ts.stress > 0.0  # the platens load the specimen
    c = g.counters()
    assert c.bonds > 4 * c.n


@pytest.mark.timeout(300)
def test_c2_uniaxial_100k_breaks_and_contacts():
    """C2's own sample (100k spheres from the benchmark's generator, stored in
    spatial order) with its platens driven at 2 m/s, so that within 2000
    steps about 300 bonds break next to the platens and unbonded pairs
    touch: every chunk bitwise against the oracle."""
    print("case C2 100k uniaxial, platens at 2 m/s", flush=True)
    w = config.config_workload("C2")
    w["cfg"] = config.uniaxial_cfg(w["xyz"], w["r"], v_load_mps=-2.0, dnsave=50)
    g, o = _run_pair(w, (1, 999, 1000), "C2")
    c = g.counters()
    counts = g.bond_counts()
    print(f"C2: bond count {counts[0]} -> {counts[-1]}, contacts {c.contacts}, "
          f"rebuilds {c.rebuilds}", flush=True)
    assert counts[-1] < counts[0]  # bonds broke
    assert c.contacts > 0          # unbonded pairs touch


@pytest.mark.timeout(300)
def test_c1_whole_test_macro_parity():
    """C1's sample and material in a whole uniaxial test through the drop-in
    entry point (rf_run_uniaxial), platens at 1 m/s, loading flagged above
    2 MPa, until the fracture criterion or 3 % strain ends it, against the
    oracle run to the same end: peak stress (UCS), peak strain, secant
    modulus, linearity deviation, destruction, step count, curve and
    bond-count trace identical, so their error is 0."""
    print("case C1 whole uniaxial test, platens at 1 m/s", flush=True)
    w = config.config_workload("C1")
    cfg = config.uniaxial_cfg(w["xyz"], w["r"], v_load_mps=-1.0, min_stress_mpa=2.0,
                              max_strain=0.03, dnsave=20, max_steps=100000)
    g = Context(w["xyz"], w["r"], w["mat"], w["box"])
    o = pyoracle.Oracle(w["xyz"], w["r"], w["mat"], w["box"], threads=THREADS)
    g.build_bonds()
    o.build_bonds()
    res = g.run_uniaxial(cfg)
    o.set_test(cfg)
    o.step(100000)
    o.post()
    t = o.test_state()
    print(f"C1: peak {res.peak_stress * 1e3:.4f} MPa at strain {res.peak_strain:.6f}, "
          f"secant {res.secant_modulus:.4f} GPa, LD {res.linearity_deviation:.6f}, "
          f"destruction {res.destruction:.6f}, fracture {res.fracture}, {res.steps} steps",
          flush=True)
    assert t.done == 1 and res.steps == t.step
    assert res.fracture == t.fracture
    assert res.peak_stress == t.peak_stress and res.peak_strain == t.peak_strain
    assert res.secant_modulus == t.peak_stress / t.peak_strain
    assert res.linearity_deviation == t.linearity_deviation
    assert res.destruction == t.destruction and res.destruction > 0.0
    np.testing.assert_array_equal(g.curve(), o.curve())
    np.testing.assert_array_equal(g.bond_counts(), o.bond_counts())
    compare_test_states(g.test_state(), t, what="end of the C1 test")


@pytest.mark.timeout(300)
def test_c3_brazilian_50k():
    print("case C3 50k Brazilian 0.02 m/s", flush=True)
    w = config.config_workload("C3")
    g, _ = _run_pair(w, (1, 99, 400), "C3")
    assert g.test_state().step == 500


@pytest.mark.timeout(600)
def test_c4_batch_64_samples():
    print("case C4 64-sample batch", flush=True)
    ws = config.batch_workloads(64)
    steps = 200

    def make(w):
        c = Context(w["xyz"], w["r"], w["mat"], w["box"])
        c.build_bonds()
        c.set_test(w["cfg"])
        return c

    bat = [make(w) for w in ws]
    b = Batch(bat)
    done = b.step(steps)
    assert all(d == steps for d in done), done
    print(f"C4: batch of {len(ws)} samples, {sum(w['n'] for w in ws)} spheres stepped", flush=True)
    singles = []
    for k, w in enumerate(ws):
        s = make(w)
        assert s.step(steps) == steps
        compare_states(bat[k].get_state(), s.get_state(), s.n, s.slots,
                       what=f"C4 sample {k} batch vs single")
        compare_test_states(bat[k].test_state(), s.test_state(), what=f"C4 sample {k}")
        np.testing.assert_array_equal(bat[k].curve(), s.curve())
        singles.append(s)
    print("C4: every sample equals its single run", flush=True)
    order = np.argsort([w["n"] for w in ws])
    for k in (int(order[0]), int(order[len(order) // 2])):
        w = ws[k]
        o = pyoracle.Oracle(w["xyz"], w["r"], w["mat"], w["box"], threads=THREADS)
        o.build_bonds()
        o.set_test(w["cfg"])
        assert o.step(steps) == steps
        compare_states(bat[k].get_state(), o.get_state(), w["n"], 32,
                       what=f"C4 sample {k} (n={w['n']}) vs oracle")
        compare_test_states(bat[k].test_state(), o.test_state(), what=f"C4 sample {k} oracle")
        np.testing.assert_array_equal(bat[k].curve(), o.curve())
        print(f"C4: sample {k} (n={w['n']}) equals the oracle", flush=True)


@pytest.mark.timeout(300)
def test_c5_gradient_window_forward():
    print("case C5 20k 2000-step window forward", flush=True)
    w = config.gradient_workload()
    g, _ = _run_pair(w, (1000, 1000), "C5")
    ts = g.test_state()
    assert ts.step == 2000 and ts.stress > 0.0

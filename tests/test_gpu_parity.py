"""GPU (HIP) path against the CPU FP64 oracle on identical seeded samples.

Bar: interaction lists, bond numbering and the bond-break sequence identical;
particle state, forces, moments and histories bitwise identical (the two
sides follow one floating-point operation order); controller state and
curves identical.
"""
import numpy as np
import pytest

import pyoracle
from helpers import compare_states, compare_test_states, randomize_motion
from rockfit import abi, config
from rockfit.core import Context

pytestmark = pytest.mark.gpu


def _pair(w, slots=32, skin=-1.0, order=abi.RF_ORDER_SPATIAL):
    g = Context(w["xyz"], w["r"], w["mat"], w["box"], slots=slots, skin=skin, order=order)
    o = pyoracle.Oracle(w["xyz"], w["r"], w["mat"], w["box"], slots=slots, threads=4, order=order)
    return g, o


@pytest.fixture(scope="module")
def small_uniaxial():
    return config.make_workload(n=2000, r_min=0.2, r_max=0.3, test="uniaxial",
                                test_kwargs=dict(v_load_mps=-2.0, dnsave=5, max_strain=0.05))


def test_build_parity(small_uniaxial):
    w = small_uniaxial
    g, o = _pair(w)
    g.build_bonds()
    o.build_bonds()
    gs, os_ = g.get_state(), o.get_state()
    compare_states(gs, os_, g.n, g.slots, what="after build")
    assert np.count_nonzero(os_["type"]) > 3 * g.n  # a bonded sample


def test_uniaxial_steps_parity(small_uniaxial):
    w = small_uniaxial
    g, o = _pair(w)
    g.build_bonds()
    o.build_bonds()
    g.set_test(w["cfg"])
    o.set_test(w["cfg"])
    done = 0
    for chunk in (1, 9, 90, 200, 300):
        kg = g.step(chunk)
        ko = o.step(chunk)
        assert kg == ko
        done += kg
        compare_states(g.get_state(), o.get_state(), g.n, g.slots, what=f"step {done}")
        compare_test_states(g.test_state(), o.test_state(), what=f"step {done}")
    curve = g.curve()
    assert np.any(curve[:, 2] != 0.0), "platens never reached the specimen"
    np.testing.assert_array_equal(curve, o.curve())
    np.testing.assert_array_equal(g.bond_counts(), o.bond_counts())


@pytest.mark.parametrize("gap_factor", [0.3, 1.0])
def test_dense_bonds_parity(gap_factor):
    """Bond gaps of 0.3 and 1.0 r_min: 37 % / 83 % of the particles hold more
    than 8 bonds (up to 14 / 19), so the step kernel's waves take every bond
    path: bonds 9-16 evaluated by lanes without a bond of their own, the
    second round where a wave has more such bonds than free lanes, and a
    third round past 16 bonds."""
    print(f"case dense bonds gap_factor={gap_factor}", flush=True)
    w = config.make_workload(n=1500, r_min=0.2, r_max=0.3, test="uniaxial", seed=31,
                             mat_kwargs=dict(gap_factor=gap_factor),
                             test_kwargs=dict(v_load_mps=-2.0, dnsave=5, max_strain=0.05))
    g, o = _pair(w)
    g.build_bonds()
    o.build_bonds()
    nb = (np.asarray(o.get_state()["type"]).reshape(o.n, -1) == 1).sum(axis=1)
    assert (nb > 8).mean() > 0.3
    if gap_factor >= 1.0:
        assert nb.max() > 16
    g.set_test(w["cfg"])
    o.set_test(w["cfg"])
    done = 0
    for chunk in (1, 49, 150):
        assert g.step(chunk) == o.step(chunk)
        done += chunk
        compare_states(g.get_state(), o.get_state(), g.n, g.slots, what=f"step {done}")
        compare_test_states(g.test_state(), o.test_state(), what=f"step {done}")
    np.testing.assert_array_equal(g.curve(), o.curve())


def test_free_motion_append_parity():
    """Random velocities, no platens: contacts open and close, rows are
    appended and pruned every step; small skin forces frequent rebuilds."""
    w = config.make_workload(n=1500, r_min=0.2, r_max=0.3, test="none",
                             mat_kwargs=dict(sigma_mpa=5.0, tau_mpa=8.0))
    g, o = _pair(w, skin=0.02)
    g.build_bonds()
    o.build_bonds()
    st = randomize_motion(o.get_state(), g.n, 7, 2e-3, 2e-2)
    g.set_state(st)
    o.set_state(st)
    cfg = abi.PlatenCfg()
    cfg.kind = abi.RF_TEST_NONE
    g.set_test(cfg)
    o.set_test(cfg)
    for chunk in (1, 20, 80):
        assert g.step(chunk) == o.step(chunk) == chunk
        compare_states(g.get_state(), o.get_state(), g.n, g.slots, what=f"free motion +{chunk}")
    c = g.counters()
    assert c.rebuilds > 2


def test_breakage_parity():
    """Weak bonds under fast loading: the break sequence must match."""
    w = config.make_workload(n=2000, r_min=0.2, r_max=0.3, test="uniaxial",
                             mat_kwargs=dict(sigma_mpa=8.0, tau_mpa=12.0),
                             test_kwargs=dict(v_load_mps=-2.0, dnsave=1, max_steps=500,
                                              max_strain=0.05))
    g, o = _pair(w)
    g.build_bonds()
    o.build_bonds()
    g.set_test(w["cfg"])
    o.set_test(w["cfg"])
    assert g.step(500) == o.step(500)
    bg, bo = g.bond_counts(), o.bond_counts()
    np.testing.assert_array_equal(bg, bo)
    assert bg[-1] < bg[0] // 2, "expected the specimen to fail"
    assert g.test_state().fracture == 1
    compare_states(g.get_state(), o.get_state(), g.n, g.slots, what="breakage")
    compare_test_states(g.test_state(), o.test_state(), what="breakage")


def test_brazilian_parity():
    w = config.make_workload(n=2000, r_min=0.2, r_max=0.3, shape="disc", aspect=0.5,
                             test="brazilian", test_kwargs=dict(v_load_mps=-2.0, dnsave=5,
                                                                max_steps=400))
    g, o = _pair(w)
    g.build_bonds()
    o.build_bonds()
    g.set_test(w["cfg"])
    o.set_test(w["cfg"])
    assert g.step(400) == o.step(400)
    compare_states(g.get_state(), o.get_state(), g.n, g.slots, what="brazilian")
    compare_test_states(g.test_state(), o.test_state(), what="brazilian")
    np.testing.assert_array_equal(g.curve(), o.curve())


def test_graph_replay_parity():
    """Full chunks replayed as captured HIP graphs (with and without timing
    events inside) give the same state as plain launches, over several
    replays with host work in between."""
    w = config.make_workload(n=1500, r_min=0.2, r_max=0.3, test="uniaxial", seed=5,
                             test_kwargs=dict(v_load_mps=-2.0, dnsave=5, max_steps=2000,
                                              max_strain=0.05))
    ctxs = []
    for graph, timing in ((True, 0), (False, 0), (True, 4)):  # timing: one sample per chunk
        c = Context(w["xyz"], w["r"], w["mat"], w["box"], skin=0.03)
        c.build_bonds()
        c.set_test(w["cfg"])
        c.set_graph(graph)
        c.set_timing(timing)
        ctxs.append(c)
    for chunk in (600, 300, 400):
        done = [c.step(chunk) for c in ctxs]
        assert done[0] == done[1] == done[2]
        ref = ctxs[1].get_state()
        for k in (0, 2):
            compare_states(ctxs[k].get_state(), ref, w["n"], 32, what=f"graph case {k}")
            compare_test_states(ctxs[k].test_state(), ctxs[1].test_state())
    kms, timed, _, _ = ctxs[2].last_timing()
    assert timed > 0 and kms > 0.0
    assert ctxs[0].counters().rebuilds > 1


def test_ejected_particles_parity():
    """Sixty surface spheres shot outward leave the space box and are parked
    at the hidden point (d_ParallelepipedCutRestriction): more hidden
    particles than a candidate list holds, all at one point.  State, curve
    and controller stay bitwise equal, and parked particles do not keep
    asking for list rebuilds."""
    w = config.make_workload(n=1500, r_min=0.2, r_max=0.3, test="none", seed=8)
    g, o = _pair(w)
    g.build_bonds()
    o.build_bonds()
    free = abi.PlatenCfg()
    free.kind = abi.RF_TEST_NONE
    g.set_test(free)
    o.set_test(free)
    st = g.get_state()
    n = g.n
    R = st["R"].reshape(3, n)
    c = R.mean(axis=1)
    rad = np.hypot(R[0] - c[0], R[1] - c[1])
    out = np.argsort(-rad, kind="stable")[:60]
    V = st["V"].reshape(3, n)
    V[0, out] = 0.8 * (R[0, out] - c[0]) / rad[out]
    V[1, out] = 0.8 * (R[1, out] - c[1]) / rad[out]
    g.set_state(st)
    o.set_state(st)
    hid = np.array(w["box"].hidden[:])
    for chunk in (50, 250, 700):
        assert g.step(chunk) == o.step(chunk) == chunk
        compare_states(g.get_state(), o.get_state(), n, g.slots, what=f"ejected +{chunk}")
    R = g.get_state()["R"].reshape(3, n)
    parked = np.all(R.T == hid, axis=1).sum()
    assert parked > 32, parked
    before = g.counters().rebuilds
    assert g.step(500) == o.step(500) == 500
    compare_states(g.get_state(), o.get_state(), n, g.slots, what="ejected, parked")
    assert g.counters().rebuilds - before < 50


@pytest.mark.parametrize("kind", ["uniaxial", "brazilian"])
def test_full_run_macro_parity(kind):
    """A whole test, loading at 0.5 m/s until the fracture criterion ends it,
    through the drop-in entry point (rf_run_uniaxial / rf_run_brazilian)
    against the oracle run to the same end: the macroscopic results (peak
    stress, i.e. UCS or tensile strength, peak strain, secant modulus,
    linearity deviation, destruction), the step count, the whole curve and
    the bond-count trace are identical, so their error is 0."""
    print(f"case full {kind} run n=1500 v=-0.5 m/s", flush=True)
    shape, aspect = ("cylinder", 2.0) if kind == "uniaxial" else ("disc", 0.5)
    w = config.make_workload(n=1500, r_min=0.2, r_max=0.3, test=kind, shape=shape, aspect=aspect,
                             seed=12, test_kwargs=dict(v_load_mps=-0.5, dnsave=20, max_steps=30000))
    g, o = _pair(w)
    g.build_bonds()
    o.build_bonds()
    res = g.run_uniaxial(w["cfg"]) if kind == "uniaxial" else g.run_brazilian(w["cfg"])
    o.set_test(w["cfg"])
    o.step(30000)
    o.post()
    t = o.test_state()
    assert t.done == 1 and t.fracture == 1 and res.fracture == 1
    print(f"{kind}: peak {res.peak_stress * 1e3:.4f} MPa at strain {res.peak_strain:.6f}, "
          f"secant {res.secant_modulus:.4f} GPa, LD {res.linearity_deviation:.6f}, "
          f"destruction {res.destruction:.6f}, {res.steps} steps", flush=True)
    assert res.peak_stress == t.peak_stress and res.peak_strain == t.peak_strain
    assert res.secant_modulus == t.peak_stress / t.peak_strain
    assert res.linearity_deviation == t.linearity_deviation
    assert res.destruction == t.destruction
    assert res.steps == t.step and res.curve_points == t.n_curve
    np.testing.assert_array_equal(g.curve(), o.curve())
    np.testing.assert_array_equal(g.bond_counts(), o.bond_counts())
    compare_test_states(g.test_state(), t, what=f"end of the {kind} run")


def test_caller_stream():
    """A context on a stream the caller made (rf_options.stream) runs the same
    steps as one on its own stream, bit for bit, and rf_destroy leaves that
    stream to the caller (it still synchronises and is destroyed here)."""
    import ctypes

    print("case caller stream n=1500", flush=True)
    hip = ctypes.CDLL("libamdhip64.so")
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    w = config.make_workload(n=1500, r_min=0.2, r_max=0.3, test="uniaxial", seed=9,
                             test_kwargs=dict(v_load_mps=-2.0, dnsave=5, max_steps=600))
    a = Context(w["xyz"], w["r"], w["mat"], w["box"], stream=stream.value)
    b = Context(w["xyz"], w["r"], w["mat"], w["box"])
    for c in (a, b):
        c.build_bonds()
        c.set_test(w["cfg"])
        assert c.step(600) == 600
    compare_states(a.get_state(), b.get_state(), a.n, a.slots, what="caller stream")
    np.testing.assert_array_equal(a.curve(), b.curve())
    a.close()
    b.close()
    assert hip.hipStreamSynchronize(stream) == 0  # still the caller's
    assert hip.hipStreamDestroy(stream) == 0


def test_full_size_parity():
    """The benchmark's workload at its full size: 100,000 spheres,
    r ~ U[0.20, 0.35] mm, 2:1 cylinder, uniaxial loading at strain rate
    0.1/s.  Build and 300 loading steps (the platens close in and start to
    load the specimen) equal the oracle's bit for bit: particle state,
    slot table, bond numbering, histories, controller and curve."""
    import os

    print("case full size n=100000 300 steps", flush=True)
    w = config.make_workload(n=100000, r_min=0.20, r_max=0.35, test="uniaxial")
    w["cfg"] = config.uniaxial_cfg(w["xyz"], w["r"], v_load_mps=-0.1 * w["cfg"].size[2] * 1e-3,
                                   dnsave=10)
    g = Context(w["xyz"], w["r"], w["mat"], w["box"])
    o = pyoracle.Oracle(w["xyz"], w["r"], w["mat"], w["box"],
                        threads=min(16, os.cpu_count() or 1))
    g.build_bonds()
    o.build_bonds()
    compare_states(g.get_state(), o.get_state(), g.n, g.slots, what="full size after build")
    g.set_test(w["cfg"])
    o.set_test(w["cfg"])
    assert g.step(300) == o.step(300) == 300
    compare_states(g.get_state(), o.get_state(), g.n, g.slots, what="full size, 300 steps")
    compare_test_states(g.test_state(), o.test_state(), what="full size, 300 steps")
    np.testing.assert_array_equal(g.curve(), o.curve())
    np.testing.assert_array_equal(g.bond_counts(), o.bond_counts())
    print(f"stress after 300 steps {g.test_state().stress * 1e3:.6f} MPa", flush=True)


def test_wide_skin_many_candidates_parity():
    """A wide candidate skin (0.35 acut): most particles keep more than 8
    candidates, so the step kernel's late loads of candidate entries 8..31
    run in most waves; free motion with weak bonds appends and prunes
    pairs.  State and lists equal the oracle's bit for bit."""
    print("case wide skin n=1500", flush=True)
    w = config.make_workload(n=1500, r_min=0.2, r_max=0.3, test="none", seed=17,
                             mat_kwargs=dict(sigma_mpa=5.0, tau_mpa=8.0))
    g, o = _pair(w, skin=0.35 * w["mat"].acut)
    g.build_bonds()
    o.build_bonds()
    c0 = g.counters()
    assert c0.candidates > 8 * c0.n, "expected more than 8 candidates per particle"
    st = randomize_motion(o.get_state(), g.n, 9, 2e-3, 2e-2)
    g.set_state(st)
    o.set_state(st)
    cfg = abi.PlatenCfg()
    cfg.kind = abi.RF_TEST_NONE
    g.set_test(cfg)
    o.set_test(cfg)
    for chunk in (1, 30, 120):
        assert g.step(chunk) == o.step(chunk) == chunk
        compare_states(g.get_state(), o.get_state(), g.n, g.slots, what=f"wide skin +{chunk}")


def test_high_slots_parity():
    """A neighbour cutoff of 3.45 mean radii: about one particle in ten
    holds partners in slots 24..31 (up to 29 partners), so waves with and
    without such particles both run (the step kernel skips those slots'
    loads in waves where no particle uses them).  Build and 400 loading
    steps equal the oracle's bit for bit."""
    print("case high slots n=1500 acut 3.45 r_mean", flush=True)
    w = config.make_workload(n=1500, r_min=0.2, r_max=0.3, test="uniaxial", seed=33,
                             mat_kwargs=dict(acut_factor=3.45),
                             test_kwargs=dict(v_load_mps=-2.0, dnsave=5, max_strain=0.05))
    g, o = _pair(w)
    g.build_bonds()
    o.build_bonds()
    nbr = np.asarray(o.get_state()["nbr"]).reshape(o.n, -1)
    hi = (nbr[:, 24:] >= 0).any(axis=1)
    assert 50 < hi.sum() < o.n // 2
    compare_states(g.get_state(), o.get_state(), g.n, g.slots, what="high slots, build")
    g.set_test(w["cfg"])
    o.set_test(w["cfg"])
    assert g.step(400) == o.step(400) == 400
    compare_states(g.get_state(), o.get_state(), g.n, g.slots, what="high slots, 400 steps")
    compare_test_states(g.test_state(), o.test_state(), what="high slots, 400 steps")
    np.testing.assert_array_equal(g.curve(), o.curve())


def test_given_order_parity():
    """The caller's particle order kept as the storage order (RF_ORDER_GIVEN)
    instead of the spatial one: random motion with appends and prunes, and a
    platen test, equal to the oracle run in the same order."""
    w = config.make_workload(n=1500, r_min=0.2, r_max=0.3, test="uniaxial", seed=13,
                             mat_kwargs=dict(sigma_mpa=5.0, tau_mpa=8.0),
                             test_kwargs=dict(v_load_mps=-2.0, dnsave=5, max_strain=0.05))
    print("case given order n=1500", flush=True)
    g, o = _pair(w, skin=0.02, order=abi.RF_ORDER_GIVEN)
    g.build_bonds()
    o.build_bonds()
    st = randomize_motion(o.get_state(), g.n, 11, 1e-3, 1e-2)
    g.set_state(st)
    o.set_state(st)
    g.set_test(w["cfg"])
    o.set_test(w["cfg"])
    for chunk in (1, 30, 120):
        assert g.step(chunk) == o.step(chunk) == chunk
        compare_states(g.get_state(), o.get_state(), g.n, g.slots, what=f"given order +{chunk}")
        compare_test_states(g.test_state(), o.test_state(), what=f"given order +{chunk}")
    np.testing.assert_array_equal(g.curve(), o.curve())

"""Pin the CPU oracle's platen drivers, escapee cut, stress bookkeeping and
bond count against closed forms.

One or a few spheres between the platens of a uniaxial (z) or Brazilian (y)
test, with the platen geometry set by hand:

- a10 platen contact: Hertz against a rigid plane (R* = r, m* = m):
  |F| = 4/3 E* sqrt(r) delta^1.5 with E* = E/(1 - nu^2) for the uniaxial
  wall moduli (2 hm_E) and half that for the Brazilian ones (hm_E); Mindlin
  K_t = 8 G* sqrt(r delta) with G* = G/(2 - nu); the tangential spring holds
  no history from one step to the next, so its force is K_t v_t dt plus the
  damping 2 sqrt(5/6) |beta| sqrt(K_t m) v_t, capped at mu |F_n|; a sphere
  bouncing off a fixed platen leaves with the restitution coefficient;
- a12 cut: a particle leaving the space box is parked at the hidden point
  with V = F = 0; one inside takes the exact kick-drift step with
  m = rho 4/3 pi r^3;
- a2 controller: platens approach at 10 V while no force acts, each moving
  V dt / 2; a platen pair with more than 20 % force imbalance above 2 MPa
  pauses; curve points are recorded only while a platen moves;
- a13 stress: (|F_bottom| + |F_top|) / 2 over pi (R + r_ref)^2 (uniaxial) or
  pi R t (Brazilian); loading flag, peak tracking and the two fracture rules;
- a14 bond count: twice the number of pairs with rho < r_i + r_j + gap.

The references are md_compression_hm.cu:20-449, md_parallelepipedrestriction.cu:
73-160, md_calculateincrementsdem.cu:19-49, md_data_types.h:212-242, 347-409,
462-494 and md_results.cu:56-145, 189-272.
"""
import math

import numpy as np
import pytest

import pyoracle
from rockfit import abi, config


R0 = 0.5


def _mat(**kw):
    kw.setdefault("gap_factor", -1.0)  # no bonds unless asked for
    return config.material(R0, R0, **kw)


def _setup(mat, offsets, radii, space=(20.0, 20.0, 20.0)):
    box = config.box_reference(list(space), 2.0 * mat.acut)
    c = np.array(box.center[:])
    xyz = c + np.asarray(offsets, float)
    o = pyoracle.Oracle(xyz, np.asarray(radii, float), mat, box)
    o.build_bonds()
    return o, c, box


def _cfg(kind, c, top, bottom, half_x=2.0, thick=4.0, v_load=0.0, min_stress=1e-3,
         cond=0.1, dnsave=1, max_strain=0.5):
    """Platens at absolute positions top/bottom along the test axis (z for
    the uniaxial test, y for the Brazilian one), specimen centred at c."""
    cfg = abi.PlatenCfg()
    cfg.kind = kind
    cfg.dnsave = dnsave
    cfg.v_load = v_load
    cfg.min_stress = min_stress
    cfg.max_strain = max_strain
    cfg.cond_fracture = cond
    ax = 1 if kind == abi.RF_TEST_BRAZILIAN else 2
    size = [2.0 * half_x, 2.0 * half_x, thick]
    size[ax] = top - bottom
    center = list(c)
    center[ax] = 0.5 * (top + bottom)
    for k in range(3):
        cfg.size[k] = size[k]
        cfg.center[k] = center[k]
    cfg.r_ref = R0
    cfg.wall_modulus_factor = 1.0 if kind == abi.RF_TEST_BRAZILIAN else 2.0
    cfg.max_steps = 0
    cfg.curve_cap = 0
    return cfg


def _E_star(mat):
    return mat.E_contact / (1.0 - mat.nu ** 2)


def _G_star(mat):
    G = mat.E_contact / (2.0 * (1.0 + mat.nu))
    return G / (2.0 - mat.nu)


def _beta(e):
    return abs(math.log(e)) / math.sqrt(math.pi ** 2 + math.log(e) ** 2)


def _mass(mat, r):
    return mat.density * 4.0 / 3.0 * math.pi * r ** 3


def _forces(o):
    st = o.get_state()
    n = o.n
    return st["F"].reshape(3, n).T.copy(), st["M"].reshape(3, n).T.copy(), st


@pytest.mark.parametrize("delta", [1e-4, 1e-3, 4e-3])
def test_uniaxial_platen_is_hertz_on_a_rigid_plane(delta):
    mat = _mat()
    o, c, _ = _setup(mat, [(0, 0, 0)], [R0])
    top = c[2] + R0 - delta
    cfg = _cfg(abi.RF_TEST_UNIAXIAL, c, top, c[2] - R0 - 1.0)
    o.set_test(cfg)
    o.call("ora_forces")
    F, M, _ = _forces(o)
    expect = 4.0 / 3.0 * _E_star(mat) * math.sqrt(R0) * delta ** 1.5
    assert F[0, 2] == pytest.approx(-expect, rel=1e-12)
    assert F[0, 0] == 0.0 and F[0, 1] == 0.0
    assert np.all(M == 0.0)
    # the bottom partial stays -0.0, the top partial is the platen's load
    pb, pt = o.partials()
    assert pb[0] == 0.0 and pt[0] == pytest.approx(expect, rel=1e-12)


@pytest.mark.parametrize("delta", [1e-4, 2e-3])
def test_brazilian_platen_uses_half_the_wall_moduli(delta):
    mat = _mat()
    o, c, _ = _setup(mat, [(0, 0, 0)], [R0])
    bottom = c[1] - R0 + delta
    cfg = _cfg(abi.RF_TEST_BRAZILIAN, c, c[1] + R0 + 1.0, bottom)
    o.set_test(cfg)
    o.call("ora_forces")
    F, _, _ = _forces(o)
    expect = 0.5 * 4.0 / 3.0 * _E_star(mat) * math.sqrt(R0) * delta ** 1.5
    assert F[0, 1] == pytest.approx(expect, rel=1e-12)
    assert F[0, 0] == 0.0 and F[0, 2] == 0.0
    pb, pt = o.partials()
    assert pt[0] == 0.0 and pb[0] == pytest.approx(-expect, rel=1e-12)


def test_normal_damping_against_the_platen():
    mat = _mat(restitution=0.6)
    o, c, _ = _setup(mat, [(0, 0, 0)], [R0])
    delta, v = 1e-3, 2e-4
    cfg = _cfg(abi.RF_TEST_UNIAXIAL, c, c[2] + R0 - delta, c[2] - R0 - 1.0)
    o.set_test(cfg)
    st = o.get_state()
    st["V"][2] = v  # moving into the top platen
    o.set_state(st)
    o.call("ora_forces")
    F, _, _ = _forces(o)
    Sn = 2.0 * _E_star(mat) * math.sqrt(R0 * delta)
    elastic = 4.0 / 3.0 * _E_star(mat) * math.sqrt(R0) * delta ** 1.5
    damping = 2.0 * math.sqrt(5.0 / 6.0) * _beta(0.6) * math.sqrt(Sn * _mass(mat, R0)) * v
    assert F[0, 2] == pytest.approx(-(elastic + damping), rel=2e-5)


def test_tangential_spring_has_no_history_and_coulomb_cap():
    """A sphere sliding along the top platen: the force is the same on every
    step (the spring restarts from zero), K_t v dt + damping, until the
    Coulomb cap mu |F_n| takes over."""
    mat = _mat(friction=0.5, rolling_friction=0.0)
    delta = 2e-3
    Fn = 4.0 / 3.0 * _E_star(mat) * math.sqrt(R0) * delta ** 1.5
    Kt = 8.0 * _G_star(mat) * math.sqrt(R0 * delta)
    m = _mass(mat, R0)
    d = pyoracle.derive(mat, R0, R0)
    for v, capped in ((1e-6, False), (1.0, True)):
        o, c, _ = _setup(mat, [(0, 0, 0)], [R0])
        cfg = _cfg(abi.RF_TEST_UNIAXIAL, c, c[2] + R0 - delta, c[2] - R0 - 1.0)
        o.set_test(cfg)
        got = []
        for _ in range(3):
            st = o.get_state()
            st["R"][:] = c  # one sphere: x, y, z blocks of length 1
            st["V"][:] = (v, 0.0, 0.0)
            st["W"][:] = 0.0
            o.set_state(st)
            o.step(1)  # a full step: append, forces, move, controller
            F, M, _ = _forces(o)
            got.append(F[0].copy())
        for g in got[1:]:
            assert np.array_equal(g, got[0])
        ft = got[0][0]
        if capped:
            assert abs(ft) == pytest.approx(0.5 * Fn, rel=1e-12)
        else:
            damp = 2.0 * math.sqrt(5.0 / 6.0) * _beta(mat.restitution) * math.sqrt(Kt * m)
            # the code's constant 1.8257 is 2 sqrt(5/6) to 2.3e-5
            assert -ft == pytest.approx(Kt * v * d.dt + damp * v, rel=5e-5)
        assert ft < 0.0  # opposes the sliding
        # the moment of the tangential force about the centre: r n x F_t
        assert M[0, 1] == pytest.approx(R0 * ft, rel=1e-12)


def test_bounce_off_a_fixed_platen_returns_the_restitution():
    mat = _mat(restitution=0.6, friction=0.0, rolling_friction=0.0)
    o, c, _ = _setup(mat, [(0, 0, 0)], [R0])
    cfg = _cfg(abi.RF_TEST_UNIAXIAL, c, c[2] + R0 + 1e-6, c[2] - R0 - 2.0, v_load=0.0,
               min_stress=1e9)
    o.set_test(cfg)
    v0 = 1e-3
    st = o.get_state()
    st["V"][2] = v0
    o.set_state(st)
    left = False
    for _ in range(5000):
        o.step(1)
        s = o.get_state()
        if s["R"][2] < c[2] - 1e-4 and s["V"][2] < 0:
            left = True
            break
    assert left
    assert -s["V"][2] / v0 == pytest.approx(0.6, abs=0.05)
    t = o.test_state()
    assert t.top == cfg.center[2] + 0.5 * cfg.size[2]  # v_load = 0: platens fixed


def test_cut_parks_escapees_and_kick_drift_is_exact():
    mat = _mat()
    space = (20.0, 20.0, 20.0)
    box = config.box_reference(list(space), 2.0 * mat.acut)
    c = np.array(box.center[:])
    half = np.array(box.half[:])
    xyz = np.array([c, c + np.array([half[0] - 1e-3, 0, 0]), c + np.array([0, 0, -half[2] + 1e-3])])
    r = np.array([R0, 0.3, 0.4])
    o = pyoracle.Oracle(xyz, r, mat, box)
    o.build_bonds()
    d = pyoracle.derive(mat, 0.3, R0)
    st = o.get_state()
    n = 3
    V = np.array([[1e-4, -2e-4, 3e-4], [1.0, 0.0, 0.0], [0.0, 0.0, -1.0]])
    F = np.array([[2e-3, 1e-3, -5e-4], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    Mo = np.array([[1e-5, 0.0, -2e-5], [0, 0, 0], [0, 0, 0]])
    st["V"][:] = V.T.reshape(-1)
    st["F"][:] = F.T.reshape(-1)
    st["M"][:] = Mo.T.reshape(-1)
    o.set_state(st)
    o.call("ora_integrate")
    s = o.get_state()
    R1 = s["R"].reshape(3, n).T
    V1 = s["V"].reshape(3, n).T
    W1 = s["W"].reshape(3, n).T
    F1 = s["F"].reshape(3, n).T
    m = _mass(mat, R0)
    inertia = 0.4 * m * R0 * R0
    v_exp = V[0] + F[0] * d.dt / m
    np.testing.assert_allclose(V1[0], v_exp, rtol=1e-13)
    np.testing.assert_allclose(R1[0], c + v_exp * d.dt, rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(W1[0], Mo[0] * d.dt / inertia, rtol=1e-13)
    hidden = np.array(box.hidden[:])
    for i in (1, 2):  # both crossed a face of the space box
        assert np.array_equal(R1[i], hidden)
        assert np.all(V1[i] == 0.0) and np.all(F1[i] == 0.0)


def test_controller_approach_speed_and_free_travel():
    mat = _mat()
    o, c, _ = _setup(mat, [(0, 0, 0)], [R0])
    v = -2e-5
    top0, bot0 = c[2] + R0 + 0.5, c[2] - R0 - 0.5
    cfg = _cfg(abi.RF_TEST_UNIAXIAL, c, top0, bot0, v_load=v, dnsave=10)
    o.set_test(cfg)
    d = pyoracle.derive(mat, R0, R0)
    k = 40
    o.step(k)
    t = o.test_state()
    size0 = top0 - bot0
    assert t.v == 10.0 * v  # no force on either platen: 10 V
    assert t.size0 == pytest.approx(size0, rel=1e-15)
    assert t.size_axis == pytest.approx(size0 + k * 10.0 * v * d.dt, rel=1e-13)
    assert t.top == pytest.approx(top0 + k * 5.0 * v * d.dt, rel=1e-13)
    assert t.bottom == pytest.approx(bot0 - k * 5.0 * v * d.dt, rel=1e-13)
    assert t.strain == pytest.approx(k * 10.0 * v * d.dt / size0, rel=1e-10)
    # curve points every dnsave steps, the strain recorded before that step's
    # move; the point of step 0 (platens still at rest, V = 0) is overwritten
    # by the next one, so point j holds the strain after (j + 1) * dnsave moves
    cur = o.curve()
    assert t.n_curve == k // 10 - 1 and len(cur) >= k // 10 - 1
    for j in range(k // 10 - 1):
        assert cur[j, 0] == pytest.approx((j + 1) * 10 * 10.0 * v * d.dt / size0, rel=1e-10)
        assert cur[j, 1] == 0.0 and cur[j, 2] == 0.0
    assert t.stress == 0.0 and t.loading == 0 and t.fracture == 0


def test_curve_is_not_advanced_while_platens_stand_still():
    mat = _mat()
    o, c, _ = _setup(mat, [(0, 0, 0)], [R0])
    cfg = _cfg(abi.RF_TEST_UNIAXIAL, c, c[2] + R0 + 0.5, c[2] - R0 - 0.5, v_load=0.0)
    o.set_test(cfg)
    o.step(25)
    t = o.test_state()
    assert t.n_curve == 0 and t.step == 25


def _pressed(mat, delta, both, kind=abi.RF_TEST_UNIAXIAL, **kw):
    o, c, _ = _setup(mat, [(0, 0, 0)], [R0])
    ax = 1 if kind == abi.RF_TEST_BRAZILIAN else 2
    top = c[ax] + R0 - delta
    bottom = c[ax] - R0 + (delta if both else -1.0)
    cfg = _cfg(kind, c, top, bottom, **kw)
    o.set_test(cfg)
    return o, c, cfg


def test_one_sided_load_above_2mpa_pauses_the_platens():
    mat = _mat()
    v = -1e-5
    # small area (half width 0.5 mm): pi (0.5 + 0.5)^2 = pi mm^2
    for delta, pause in ((1e-5, False), (5e-3, True)):
        o, c, cfg = _pressed(mat, delta, both=False, v_load=v, half_x=0.5, min_stress=1.0)
        o.step(1)
        t = o.test_state()
        F = 4.0 / 3.0 * _E_star(mat) * math.sqrt(R0) * delta ** 1.5
        area = math.pi * (0.5 + R0) ** 2
        assert t.sfl1 == pytest.approx(F, rel=1e-12) and t.sfl0 == 0.0
        assert t.stress == pytest.approx(0.5 * F / area, rel=1e-12)
        assert (0.5 * F / area > 2e-3) == pause
        if pause:
            assert t.v == 0.0 and t.steps_stop == 1 and t.size_axis == t.size0
        else:
            assert t.v == v and t.steps_stop == 0


def test_balanced_load_keeps_loading():
    mat = _mat()
    v = -1e-5
    o, c, cfg = _pressed(mat, 3e-3, both=True, v_load=v, half_x=0.5, min_stress=1.0)
    o.step(1)
    t = o.test_state()
    F = 4.0 / 3.0 * _E_star(mat) * math.sqrt(R0) * 3e-3 ** 1.5
    assert t.sfl0 == pytest.approx(-F, rel=1e-12) and t.sfl1 == pytest.approx(F, rel=1e-12)
    area = math.pi * (0.5 + R0) ** 2
    assert t.stress == pytest.approx(F / area, rel=1e-12) and t.stress > 2e-3
    assert t.v == v and t.steps_stop == 0


@pytest.mark.parametrize("kind", [abi.RF_TEST_UNIAXIAL, abi.RF_TEST_BRAZILIAN])
def test_loading_peak_and_fracture_when_the_load_vanishes(kind):
    mat = _mat()
    delta = 6e-3
    min_stress = 1e-3
    o, c, cfg = _pressed(mat, delta, both=True, kind=kind, v_load=0.0, half_x=1.0,
                         thick=2.0, min_stress=min_stress)
    o.step(1)
    t = o.test_state()
    factor = 0.5 if kind == abi.RF_TEST_BRAZILIAN else 1.0
    F = factor * 4.0 / 3.0 * _E_star(mat) * math.sqrt(R0) * delta ** 1.5
    area = math.pi * 1.0 * 2.0 if kind == abi.RF_TEST_BRAZILIAN else math.pi * (1.0 + R0) ** 2
    assert t.stress == pytest.approx(F / area, rel=1e-12)
    assert t.stress > min_stress and t.loading == 1 and t.fracture == 0
    assert t.peak_stress == t.stress and t.peak_strain == 0.0
    # the specimen falls apart: no load on the next step
    st = o.get_state()
    st["R"][0] += 3.0  # out of the platens' scope, still inside the space box
    st["V"][:] = 0.0
    o.set_state(st)
    o.step(1)
    t2 = o.test_state()
    assert t2.stress == 0.0 and t2.fracture == 1 and t2.done == 1
    if kind == abi.RF_TEST_BRAZILIAN:
        # the Brazilian peak follows a drop below cond * peak
        assert t2.peak_stress == 0.0
    else:
        assert t2.peak_stress == t.stress


def test_uniaxial_fracture_below_cond_times_peak():
    """Peak at a one-sided load, then loads of 0.55 and 0.45 of it: only
    the second is below cond_fracture = 0.5 of the peak."""
    mat = _mat()
    d1 = 6e-3
    o, c, cfg = _pressed(mat, d1, both=False, v_load=0.0, half_x=1.0, min_stress=1e-4,
                         cond=0.5)
    o.step(1)
    t = o.test_state()
    assert t.loading == 1 and t.fracture == 0
    for ratio, broken in ((0.55, 0), (0.45, 1)):
        st = o.get_state()
        st["R"][2] = c[2] - d1 * (1.0 - ratio ** (2.0 / 3.0))  # top overlap d1 ratio^(2/3)
        st["V"][:] = 0.0
        st["W"][:] = 0.0
        o.set_state(st)
        o.step(1)
        t2 = o.test_state()
        assert t2.stress / t.peak_stress == pytest.approx(ratio, rel=1e-9)
        assert t2.stress > 0.33 * cfg.min_stress
        assert t2.fracture == broken and t2.peak_stress == t.peak_stress


def test_bond_count_is_twice_the_bonded_pairs():
    rng = np.random.default_rng(7)
    mat = config.material(0.2, 0.25, gap_factor=0.3)
    # a jittered cubic lattice, spacing 0.45: about half of the lattice
    # neighbours fall under r_i + r_j + gap, diagonals never do
    g = np.arange(4) * 0.45 - 0.675
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pts = pts + rng.uniform(-0.02, 0.02, pts.shape)
    n = len(pts)
    r = rng.uniform(0.19, 0.22, n)
    o, c, box = _setup(mat, pts, r)
    xyz = np.asarray(pts)
    gap = mat.bond_gap
    pairs = 0
    for i in range(n):
        for j in range(i + 1, n):
            rho = np.linalg.norm(xyz[i] - xyz[j])
            if rho < r[i] + r[j] + gap:
                pairs += 1
    assert 40 < pairs < 144  # 144 lattice neighbour pairs
    cfg = _cfg(abi.RF_TEST_UNIAXIAL, c, c[2] + 5.0, c[2] - 5.0, v_load=-1e-5, dnsave=1)
    o.set_test(cfg)
    o.step(2)  # step 0's sample is overwritten by step 1's (platens at rest at step 0)
    assert len(o.bond_counts()) == 1 and o.bond_counts()[0] == 2 * pairs
    assert o.counters().bonds == 2 * pairs

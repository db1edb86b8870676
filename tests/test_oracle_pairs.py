"""Pin the CPU oracle's pair physics against textbook closed forms.

Two-sphere set-ups in an empty box: Hertz normal force, Mindlin tangential
stiffness, Coulomb cap, rolling friction, parallel-bond axial / shear /
torsion / bending stiffness, breakage threshold, force antisymmetry, the
bonded-pair oscillation period (integrator + bond law) and the head-on
restitution (damping constant 1.8257 = 2 sqrt(5/6)).
"""
import math

import numpy as np
import pytest

import pyoracle
from rockfit import abi, config


def _mat(**kw):
    m = config.material(0.5, 0.5, **kw)
    return m


def _two(mat, p0, p1, r0=0.5, r1=0.5, space=(20.0, 20.0, 20.0)):
    box = config.box_reference(list(space), 2.0 * mat.acut)
    c = np.array(box.center[:])
    xyz = np.array([c + np.asarray(p0, float), c + np.asarray(p1, float)])
    r = np.array([r0, r1], float)
    return pyoracle.Oracle(xyz, r, mat, box), c


def _forces(o):
    st = o.get_state()
    n = o.n
    F = st["F"].reshape(3, n).T.copy()
    M = st["M"].reshape(3, n).T.copy()
    return F, M, st


def _build_no_bond(o):
    o.call("ora_construct_rows")
    o.call("ora_construct_bonds")
    o.call("ora_assign_strengths")
    o.call("ora_prune")


def test_hertz_normal_force():
    mat = _mat(gap_factor=-1.0)  # negative gap: no bonds
    d = ora_der = None
    for delta in (1e-4, 1e-3, 5e-3):
        o, _ = _two(mat, (0, 0, 0), (1.0 - delta, 0, 0), 0.5, 0.5)
        _build_no_bond(o)
        o.call("ora_forces")
        F, _, _ = _forces(o)
        d = pyoracle.derive(mat, 0.5, 0.5)
        E_star = mat.E_contact / (2.0 * (1.0 - mat.nu ** 2))  # identical spheres
        R_star = 0.25
        expect = 4.0 / 3.0 * E_star * math.sqrt(R_star) * delta ** 1.5
        assert F[0, 0] == pytest.approx(-expect, rel=1e-12)
        assert F[1, 0] == pytest.approx(expect, rel=1e-12)
        assert F[0, 1] == 0.0 and F[0, 2] == 0.0
    assert ora_der is None and d is not None


def test_hertz_unequal_radii():
    mat = _mat(gap_factor=-1.0)
    r0, r1, delta = 0.3, 0.45, 2e-3
    o, _ = _two(mat, (0, 0, 0), (0, r0 + r1 - delta, 0), r0, r1)
    _build_no_bond(o)
    o.call("ora_forces")
    F, _, _ = _forces(o)
    E_star = mat.E_contact / (2.0 * (1.0 - mat.nu ** 2))
    R_star = r0 * r1 / (r0 + r1)
    expect = 4.0 / 3.0 * E_star * math.sqrt(R_star) * delta ** 1.5
    assert F[1, 1] == pytest.approx(expect, rel=1e-12)


def test_mindlin_tangential_and_coulomb_cap():
    mat = _mat(gap_factor=-1.0, friction=0.5)
    delta = 2e-3
    o, _ = _two(mat, (0, 0, 0), (1.0 - delta, 0, 0))
    _build_no_bond(o)
    d = pyoracle.derive(mat, 0.5, 0.5)
    st = o.get_state()
    vt = 1e-6  # small sliding speed along y: elastic (below the cap)
    st["V"][o.n + 1] = vt  # particle 1, y component
    o.set_state(st)
    o.call("ora_forces")
    F, _, st2 = _forces(o)
    G_star = d.G_contact / (2.0 * (2.0 - mat.nu))
    Kt = 8.0 * G_star * math.sqrt(0.25 * delta)
    m = st and None
    mass = mat.density * 4.0 / 3.0 * math.pi * 0.5 ** 3
    meff = mass / 2.0
    damp = -1.8257 * d.p_A * math.sqrt(Kt * meff)
    expect_ft = Kt * vt * d.dt + damp * vt
    # force on particle 0 along +y (partner moves +y)
    assert F[0, 1] == pytest.approx(expect_ft, rel=1e-10)
    # Coulomb cap: a large sliding speed saturates |Ft| at mu |Fn|
    st["V"][o.n + 1] = 10.0
    o.set_state(st)
    o.call("ora_forces")
    F, _, _ = _forces(o)
    E_star = mat.E_contact / (2.0 * (1.0 - mat.nu ** 2))
    Fn = 4.0 / 3.0 * E_star * math.sqrt(0.25) * delta ** 1.5
    # normal damping is zero (no normal relative velocity)
    assert abs(F[0, 1]) == pytest.approx(mat.friction * Fn, rel=1e-12)
    assert m is None


def test_rolling_friction_and_roll_reset_quirk():
    delta = 2e-3
    for quirks in (abi.RF_QUIRKS_REFERENCE, 0):
        mat = _mat(gap_factor=-1.0, quirks=quirks, rolling_friction=0.05)
        o, _ = _two(mat, (0, 0, 0), (1.0 - delta, 0, 0))
        _build_no_bond(o)
        st = o.get_state()
        st["W"][0] = 1e-3  # particle 0 spins about the contact normal (x)
        o.set_state(st)
        o.call("ora_forces")
        _, M, _ = _forces(o)
        E_star = mat.E_contact / (2.0 * (1.0 - mat.nu ** 2))
        Fn = 4.0 / 3.0 * E_star * math.sqrt(0.25) * delta ** 1.5
        # rolling torque mu_r |Fn| opposing the spin (no radius factor, as the
        # live reference kernel); a spin about the normal makes no sliding
        assert M[0, 0] == pytest.approx(-mat.rolling_friction * Fn, rel=1e-9)
        # particle 1 does not spin: with the reset quirk its contact moment
        # sum is cleared; without it only the (zero) tangential torque stays
        assert M[1, 0] == 0.0 and M[1, 1] == 0.0
        assert abs(M[1, 2]) < 1e-15


def _bonded_pair(E_bond_gpa=20.0, b_ratio=1.0, sigma_mpa=1e6, tau_mpa=1e6, quirks=None, r=0.5):
    kw = dict(E_bond_gpa=E_bond_gpa, bond_radius_ratio=b_ratio, sigma_mpa=sigma_mpa,
              tau_mpa=tau_mpa, cv=0.0, gap_factor=0.05)
    if quirks is not None:
        kw["quirks"] = quirks
    mat = _mat(**kw)
    L0 = 2 * r
    o, c = _two(mat, (0, 0, 0), (L0, 0, 0), r, r)
    o.build_bonds()
    st = o.get_state()
    assert st["type"][0] == 1 and st["type"][o.S] == 1
    return o, mat, L0, c


def test_bond_axial_stiffness_and_antisymmetry():
    o, mat, L0, c = _bonded_pair()
    eps = 1e-4
    st = o.get_state()
    st["R"][1] = c[0] + L0 * (1 + eps)
    o.set_state(st)
    o.call("ora_forces")
    F, _, _ = _forces(o)
    b = 0.5
    expect = mat.E_bond * math.pi * b * b * eps  # F = E A dL / L0
    assert F[0, 0] == pytest.approx(expect, rel=1e-9)
    assert F[1, 0] == -F[0, 0]


def _bond_case(quirks, V=None, W=None):
    o, mat, L0, c = _bonded_pair(quirks=quirks)
    st = o.get_state()
    if V is not None:
        for k in range(3):
            st["V"][k * o.n + 1] = V[k]
    if W is not None:
        for k in range(3):
            st["W"][k * o.n + 1] = W[k]
    o.set_state(st)
    o.call("ora_forces")
    F, M, _ = _forces(o)
    d = pyoracle.derive(mat, 0.5, 0.5)
    return F, M, mat, d, L0


def test_bond_shear_torsion_bending():
    b = 0.5
    A = math.pi * b * b
    for quirks, a1 in ((abi.RF_QUIRKS_REFERENCE, True), (0, False)):
        # shear: partner slides along z; first step o_t = v dt (prev normal 0)
        v = 1e-4
        F, M, mat, d, L0 = _bond_case(quirks, V=(0, 0, v))
        ks = d.G_bond * A / L0
        assert F[0, 2] == pytest.approx(ks * v * d.dt, rel=1e-9)
        # asymmetric moment 0.5 (r x o_t) G A / L0: (r x o)_y = -L0 o_z
        assert M[0, 1] == pytest.approx(-0.5 * L0 * v * d.dt * ks, rel=1e-9)
        # torsion: T = G J theta / L0, J = pi b^4 / 2, theta = w dt
        w = 2e-3
        F, M, mat, d, L0 = _bond_case(quirks, W=(w, 0, 0))
        J = math.pi * b ** 4 / 2.0
        assert M[0, 0] == pytest.approx(d.G_bond * J * w * d.dt / L0, rel=1e-9)
        assert F[0, 2] == 0.0 and F[0, 1] == 0.0
        # bending about y: M = E I theta / L0, I = pi b^4 / 4; the reference
        # scales the applied moment once more by dt E / (4 L0) (A1).  The spin
        # also slides the bond ends: o_z = w L0 dt / 2 (half-sum spin x r)
        F, M, mat, d, L0 = _bond_case(quirks, W=(0, w, 0))
        I = math.pi * b ** 4 / 4.0
        bend = mat.E_bond * I * w * d.dt / L0
        if a1:
            bend *= d.dt * mat.E_bond / (4.0 * L0)
        oz = 0.5 * w * L0 * d.dt
        assert F[0, 2] == pytest.approx(ks * oz, rel=1e-9)
        assert M[0, 1] == pytest.approx(bend - 0.5 * L0 * oz * ks, rel=1e-9)


def test_bond_breaks_at_threshold():
    sig = 30.0  # MPa
    for eps, broken in ((0.9, False), (1.1, True)):
        o, mat, L0, c = _bonded_pair(E_bond_gpa=20.0, sigma_mpa=sig, tau_mpa=1e6)
        strain = eps * (sig * 1e6 / 20e9)
        st = o.get_state()
        st["R"][1] = c[0] + L0 * (1 + strain)
        o.set_state(st)
        o.call("ora_forces")
        st2 = o.get_state()
        assert (st2["type"][0] == 0) == broken
        assert (st2["type"][o.S] == 0) == broken


def test_bonded_pair_oscillation_period():
    """Undamped bond spring: period 2 pi sqrt(m_eff / k), k = E A / L0."""
    o, mat, L0, c = _bonded_pair()
    d = pyoracle.derive(mat, 0.5, 0.5)
    st = o.get_state()
    amp = 1e-5
    st["R"][1] = c[0] + L0 + amp
    o.set_state(st)
    cfg = abi.PlatenCfg()
    cfg.kind = abi.RF_TEST_NONE
    o.set_test(cfg)
    xs = []
    for _ in range(400):
        o.step(1)
        s = o.get_state()
        xs.append(s["R"][1] - s["R"][0] - L0)
    xs = np.array(xs)
    # zero crossings downward
    cross = np.nonzero((xs[:-1] > 0) & (xs[1:] <= 0))[0]
    assert len(cross) >= 2
    period = (cross[-1] - cross[0]) / (len(cross) - 1) * d.dt
    mass = mat.density * 4.0 / 3.0 * math.pi * 0.5 ** 3
    k = mat.E_bond * math.pi * 0.25 / L0
    expect = 2 * math.pi * math.sqrt((mass / 2) / k)
    assert period == pytest.approx(expect, rel=0.03)
    assert np.max(np.abs(xs)) < 1.05 * amp  # symplectic: amplitude kept


def test_head_on_restitution():
    mat = _mat(gap_factor=-1.0, restitution=0.6, friction=0.0, rolling_friction=0.0)
    o, _ = _two(mat, (0, 0, 0), (1.0 + 1e-6, 0, 0))
    _build_no_bond(o)
    st = o.get_state()
    v0 = 1e-3
    st["V"][0] = v0
    st["V"][1] = -v0
    o.set_state(st)
    cfg = abi.PlatenCfg()
    cfg.kind = abi.RF_TEST_NONE
    o.set_test(cfg)
    for _ in range(3000):
        o.step(1)
        s = o.get_state()
        gap = s["R"][1] - s["R"][0] - 1.0
        if gap > 1e-4 and s["V"][1] > 0:
            break
    e = (s["V"][1] - s["V"][0]) / (2 * v0)
    assert e == pytest.approx(0.6, abs=0.05)

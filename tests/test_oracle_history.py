"""Pin the oracle's history transport: the bond-history rotation M(t, phi)
and the contact tangential-history projection and rescale.

Reference: md_calculateforcesdem3.cu:57-101 (M), :113-118, :138-145 and
:183-186 (M applied to m_n, m_t and o_t), :264-279 (contact projection and
rescale).

The reference's 3x3 M (rows M1, M2, M3 at md_calculateforcesdem3.cu:83-91),
with t = n_prev x n and phi = n ((W_i + W_j).n) dt/2, is in vector form

    M h = h + (t - phi) x h + (t.phi) h - phi (t.h)

to first order the rotation by t - phi: the history turns with the bond
(t) and against the pair's mean spin about the bond axis (phi).  Expected
values below come from this form and from closed forms of whole rotations
(rigid turning of the bond, pure twist, turning contact normal); none
restates the oracle's nine entries.  tests/test_oracle_mutants.py checks that
a sign flip in any single entry, a transposed M, or a dropped projection or
rescale fails this module.
"""
import math

import numpy as np
import pytest

import pyoracle
from rockfit import abi, config

NO_ROLL_RESET = abi.RF_QUIRK_BEND_DOUBLE_SCALE | abi.RF_QUIRK_SLOT0_SENTINEL


def _unit(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


def _rot(axis, ang, v):
    """Rodrigues rotation of v about the unit axis by ang."""
    axis = _unit(axis)
    v = np.asarray(v, dtype=np.float64)
    return (v * math.cos(ang) + np.cross(axis, v) * math.sin(ang)
            + axis * np.dot(axis, v) * (1.0 - math.cos(ang)))


def _M_apply(t, phi, h):
    return h + np.cross(t - phi, h) + np.dot(t, phi) * h - phi * np.dot(t, h)


def _rel(got, exp):
    got = np.asarray(got, dtype=np.float64)
    exp = np.asarray(exp, dtype=np.float64)
    return float(np.max(np.abs(got - exp)) / max(np.max(np.abs(exp)), 1e-300))


def _pair(mat, p0, p1, r=0.5):
    box = config.box_reference([20.0, 20.0, 20.0], 2.0 * mat.acut)
    c = np.array(box.center[:])
    xyz = np.array([c + np.asarray(p0, float), c + np.asarray(p1, float)])
    return pyoracle.Oracle(xyz, np.array([r, r]), mat, box), c


def _vec(st, field, q):
    return st[field][3 * q:3 * q + 3]


def _set_vec(st, field, q, v):
    st[field][3 * q:3 * q + 3] = v


def _particle(st, field, i, n=2):
    return np.array([st[field][k * n + i] for k in range(3)])


def _set_particle(st, field, i, v, n=2):
    for k in range(3):
        st[field][k * n + i] = v[k]


def _bonded(n0, quirks=abi.RF_QUIRKS_REFERENCE, r=0.5):
    """A bonded pair of radius-r spheres touching along the unit vector n0
    (L0 = 2r), with strengths no load reaches."""
    mat = config.material(r, r, E_bond_gpa=20.0, bond_radius_ratio=1.0, sigma_mpa=1e12,
                          tau_mpa=1e12, cv=0.0, gap_factor=0.05, quirks=quirks)
    L0 = 2.0 * r
    o, c = _pair(mat, -0.5 * L0 * n0, 0.5 * L0 * n0, r)
    o.build_bonds()
    st = o.get_state()
    S = o.S
    k0 = int(np.nonzero(st["nbr"][:S] == 1)[0][0])
    k1 = int(np.nonzero(st["nbr"][S:2 * S] == 0)[0][0])
    assert st["type"][k0] == 1 and st["type"][S + k1] == 1
    d = pyoracle.derive(mat, r, r)
    return o, mat, d, c, L0, (k0, S + k1)


def _bond_expect(mat, d, r_vec, p_prev, Vi, Vj, Wi, Wj, hist, lam, b):
    """One bond step seen from particle i (r_vec = R_j - R_i), from the
    vector form of M and the parallel-bond increments."""
    dt, E, G = d.dt, mat.E_bond, d.G_bond
    rho = np.linalg.norm(r_vec)
    n = r_vec / rho
    t = np.cross(p_prev, n)
    phi = n * (np.dot(Wi + Wj, n) * dt * 0.5)
    dw = Wj - Wi
    mn = _M_apply(t, phi, hist["m_n"]) + n * (np.dot(n, dw) * dt * 0.5 * G * lam)
    c2 = dt * 0.25 * E * lam
    mt = _M_apply(t, phi, hist["m_t"]) + (dw - n * np.dot(n, dw)) * c2
    v = Vj - Vi - np.cross(0.5 * (Wi + Wj), r_vec)
    ot = _M_apply(t, phi, hist["o_t"]) + (v - n * np.dot(n, v)) * dt
    A = math.pi * b * b
    J4 = math.pi * b ** 4
    F = n * (E * (rho - 1.0 / lam) * lam * A) + ot * (G * lam * A)
    M = mn * J4 + mt * (c2 * J4) + 0.5 * np.cross(r_vec, ot) * (G * lam * A)
    return dict(prev_n=n, o_t=ot, m_n=mn, m_t=mt, F=F, M=M)


def test_bond_one_step_general_rotation():
    """One step with a large turn of the bond (|t| ~ 0.3) and a large mean
    spin about it (|phi| ~ 0.2), general velocities, spins and stored
    histories on both directed copies: every entry of M, its second-order
    part -phi (t.h) included, multiplies a non-zero history component."""
    rng = np.random.default_rng(5)
    n0 = _unit([0.48, -0.61, 0.63])
    o, mat, d, c, L0, (q0, q1) = _bonded(n0)
    dt = d.dt
    st = o.get_state()
    L = 1.01 * L0
    R0, R1 = c - 0.5 * L * n0, c + 0.5 * L * n0
    _set_particle(st, "R", 0, R0)
    _set_particle(st, "R", 1, R1)
    W0 = rng.normal(0.0, 0.25, 3) / dt + n0 * (0.15 / dt)
    W1 = rng.normal(0.0, 0.25, 3) / dt + n0 * (0.15 / dt)
    V0 = rng.normal(0.0, 1e-3, 3) * L0 / dt
    V1 = rng.normal(0.0, 1e-3, 3) * L0 / dt
    for i, (V, W) in enumerate(((V0, W0), (V1, W1))):
        _set_particle(st, "V", i, V)
        _set_particle(st, "W", i, W)
    hists = {}
    for q, sgn in ((q0, 1.0), (q1, -1.0)):
        p_prev = _unit(sgn * n0 + 0.3 * _unit(np.cross(n0, rng.normal(size=3))))
        h = dict(o_t=rng.normal(0, 1e-3, 3), m_n=rng.normal(0, 2e-3, 3),
                 m_t=rng.normal(0, 2e-3, 3))
        _set_vec(st, "prev_n", q, p_prev)
        for f, v in h.items():
            _set_vec(st, f, q, v)
        hists[q] = (p_prev, h)
    o.set_state(st)
    o.call("ora_forces")
    got = o.get_state()
    lam = st["inv_L0"][q0]
    assert lam == pytest.approx(1.0 / L0, rel=1e-15)
    cases = ((q0, 0, R1 - R0, V0, V1, W0, W1), (q1, 1, R0 - R1, V1, V0, W1, W0))
    for q, i, r_vec, Vi, Vj, Wi, Wj in cases:
        p_prev, h = hists[q]
        t = np.cross(p_prev, _unit(r_vec))
        phi_mag = abs(np.dot(Wi + Wj, _unit(r_vec))) * dt * 0.5
        assert np.linalg.norm(t) > 0.2 and phi_mag > 0.1
        e = _bond_expect(mat, d, r_vec, p_prev, Vi, Vj, Wi, Wj, h, lam, 0.5)
        for f in ("prev_n", "o_t", "m_n", "m_t"):
            assert _rel(_vec(got, f, q), e[f]) < 1e-12, (f, q)
        assert _rel(_particle(got, "F", i), e["F"]) < 1e-12
        assert _rel(_particle(got, "M", i), e["M"]) < 1e-12


def _preloaded_history(n0, axis, scale):
    """Shear (o_t) and bending (m_t) histories perpendicular to the bond with
    components along `axis` and along axis x n0; torsion (m_n) along it."""
    u = np.cross(axis, n0)
    return dict(o_t=scale * (0.6 * axis + 0.8 * u), m_t=scale * (-1.3 * axis + 0.5 * u),
                m_n=scale * 0.9 * n0)


def test_bond_history_corotates_under_rigid_rotation():
    """A bonded pair pre-loaded in shear, bending and torsion turns rigidly
    (W_i = W_j = w perpendicular to the bond, V = w x (R - centre)) through
    300 steps of angle x = |w| dt = 0.01: then t = sin(x) w/|w|, phi = 0,
    no increments, and M^k h is the rotation about w/|w| by k atan(sin x)
    with the perpendicular part scaled by (1 + sin^2 x)^(k/2).  In the frame
    turning with the bond the histories and the pair force are constant to
    O(k x^2)."""
    n0 = _unit([0.48, -0.61, 0.63])
    axis = _unit(np.cross(n0, [0.3, 0.9, -0.2]))
    assert np.all(np.abs(axis) > 0.1) and abs(np.dot(axis, n0)) < 1e-15
    o, mat, d, c, L0, (q0, q1) = _bonded(n0)
    dt = d.dt
    x = 0.01
    w = axis * (x / dt)
    st = o.get_state()
    L = 0.999 * L0  # also a small axial preload
    h0 = {q0: _preloaded_history(n0, axis, 1e-3), q1: _preloaded_history(-n0, axis, 1e-3)}
    for q, sgn in ((q0, 1.0), (q1, -1.0)):
        _set_vec(st, "prev_n", q, sgn * n0)
        for f, v in h0[q].items():
            _set_vec(st, f, q, v)
    nsteps = 300
    F_first = None
    for k in range(1, nsteps + 1):
        nk = _rot(axis, k * x, n0)
        for i, sgn in ((0, -1.0), (1, 1.0)):
            Ri = c + sgn * 0.5 * L * nk
            _set_particle(st, "R", i, Ri)
            _set_particle(st, "V", i, np.cross(w, Ri - c))
            _set_particle(st, "W", i, w)
        o.set_state(st)
        o.call("ora_forces")
        st = o.get_state()
        Fk = _rot(axis, -k * x, _particle(st, "F", 0))
        if F_first is None:
            F_first = Fk
    alpha = math.atan(math.sin(x))
    grow = (1.0 + math.sin(x) ** 2) ** (nsteps / 2.0)
    nN = _rot(axis, nsteps * x, n0)
    for q, sgn in ((q0, 1.0), (q1, -1.0)):
        assert _rel(_vec(st, "prev_n", q), sgn * nN) < 1e-13
        for f in ("o_t", "m_n", "m_t"):
            h = h0[q][f]
            par = np.dot(h, axis) * axis
            perp = h - par
            exp = par + grow * (math.cos(nsteps * alpha) * perp
                                + math.sin(nsteps * alpha) * np.cross(axis, perp))
            got = _vec(st, f, q)
            assert _rel(got, exp) < 1e-10, (f, q)
            # co-rotation with the bond, to O(k x^2)
            back = _rot(axis, -nsteps * x, got)
            assert _rel(back, h) < nsteps * x * x, (f, q)
    assert _rel(Fk, F_first) < nsteps * x * x
    assert np.linalg.norm(F_first) > 0.0


def test_bond_history_turns_against_twist():
    """A pair spinning rigidly about its bond axis (W_i = W_j = s n, fixed
    positions, V = 0): t = 0, phi = s dt n, no increments, so the shear and
    bending histories turn about n by -atan(s dt) per step, their lengths
    growing by (1 + (s dt)^2)^(1/2), and the torsion history along n is
    unchanged.  The sense (against the spin) is the reference's: its M turns
    the history by t - phi (md_calculateforcesdem3.cu:83-91)."""
    n0 = _unit([-0.35, 0.74, 0.57])
    axis = _unit(np.cross(n0, [0.2, 0.1, 0.9]))
    o, mat, d, c, L0, (q0, q1) = _bonded(n0)
    dt = d.dt
    sdt = 0.02
    st = o.get_state()
    for i in range(2):
        _set_particle(st, "W", i, n0 * (sdt / dt))
    h0 = {q0: _preloaded_history(n0, axis, 1e-3), q1: _preloaded_history(-n0, axis, 2e-3)}
    for q, sgn in ((q0, 1.0), (q1, -1.0)):
        _set_vec(st, "prev_n", q, sgn * n0)
        for f, v in h0[q].items():
            _set_vec(st, f, q, v)
    o.set_state(st)
    nsteps = 150
    for _ in range(nsteps):
        o.call("ora_forces")
    st = o.get_state()
    alpha = -nsteps * math.atan(sdt)
    grow = (1.0 + sdt * sdt) ** (nsteps / 2.0)
    # phi = n ((W_i + W_j).n) dt/2 is the same for both directed copies
    # (it does not change sign with n): both turn about +n0
    for q in (q0, q1):
        for f in ("o_t", "m_t", "m_n"):
            h = h0[q][f]
            par = np.dot(h, n0) * n0
            perp = h - par
            exp = par + grow * (math.cos(alpha) * perp + math.sin(alpha) * np.cross(n0, perp))
            assert _rel(_vec(st, f, q), exp) < 1e-10, (f, q)


def _contact(n0, quirks=NO_ROLL_RESET, r=0.5, delta=2e-3):
    mat = config.material(r, r, gap_factor=-1.0, quirks=quirks, friction=0.5,
                          rolling_friction=0.05)
    rho = 2.0 * r - delta
    o, c = _pair(mat, -0.5 * rho * n0, 0.5 * rho * n0, r)
    o.call("ora_construct_rows")
    o.call("ora_construct_bonds")
    o.call("ora_assign_strengths")
    o.call("ora_prune")
    st = o.get_state()
    k0 = int(np.nonzero(st["nbr"][:o.S] == 1)[0][0])
    assert st["type"][k0] == 0
    d = pyoracle.derive(mat, r, r)
    R_star = 0.5 * r
    Kn = 2.0 * d.hm_E * math.sqrt(R_star * delta)
    Kt = 8.0 * d.hm_G * math.sqrt(R_star * delta)
    fn = -delta * Kn * 2.0 / 3.0
    return o, mat, d, c, rho, k0, Kt, fn


def test_contact_history_projected_and_rescaled():
    """A stored tangential overlap that is not in the contact's tangent
    plane, no relative motion: the new overlap is the projection onto the
    plane rescaled to the old length, |o| P o / |P o| with P = I - n n^T,
    and the tangential force is K_t times it (below the Coulomb cap), with
    the moment (n x F_t) r_i."""
    n0 = _unit([0.27, 0.86, -0.43])
    o, mat, d, c, rho, k0, Kt, fn = _contact(n0)
    st = o.get_state()
    ohat = _unit([0.9, -0.2, 0.55])  # 0.38 of it along n0
    osz = 0.2 * mat.friction * abs(fn) / Kt
    ostore = ohat * osz
    _set_vec(st, "o_t", k0, ostore)
    o.set_state(st)
    o.call("ora_forces")
    got = o.get_state()
    P = ostore - n0 * np.dot(n0, ostore)
    onew = np.linalg.norm(ostore) * P / np.linalg.norm(P)
    assert _rel(_vec(got, "o_t", k0), onew) < 1e-13
    Ft = Kt * onew
    assert np.linalg.norm(Ft) < mat.friction * abs(fn)
    assert _rel(_particle(got, "F", 0), n0 * fn + Ft) < 1e-12
    assert _rel(_particle(got, "M", 0), np.cross(n0, Ft) * 0.5) < 1e-12
    # past the Coulomb cap: F_t = mu |F_n| along the rescaled overlap, and
    # the stored overlap is F_t / K_t
    st = o.get_state()
    _set_vec(st, "o_t", k0, ohat * (5.0 * mat.friction * abs(fn) / Kt))
    o.set_state(st)
    o.call("ora_forces")
    got = o.get_state()
    Pu = _unit(ohat - n0 * np.dot(n0, ohat))
    Ftc = mat.friction * abs(fn) * Pu
    assert _rel(_particle(got, "F", 0), n0 * fn + Ftc) < 1e-12
    assert _rel(_vec(got, "o_t", k0), Ftc / Kt) < 1e-12


def test_contact_history_follows_turning_normal():
    """A loaded contact whose normal turns by x = 0.05 per step about a
    fixed axis w perpendicular to it (positions rotated, no relative
    velocity): projection then rescale keep the overlap in the tangent plane
    at constant length, and with o_0 = a w + b (w x n_0) the k-th overlap is
    |o_0| (a w + b cos^k(x) (w x n_k)) / (a^2 + b^2 cos^2k(x))^(1/2)."""
    n0 = _unit([0.62, -0.31, 0.72])
    axis = _unit(np.cross(n0, [0.1, 0.95, 0.3]))
    o, mat, d, c, rho, k0, Kt, fn = _contact(n0)
    st = o.get_state()
    a, b = 0.6, 0.8
    osz = 0.3 * mat.friction * abs(fn) / Kt
    o0 = osz * (a * axis + b * np.cross(axis, n0))
    _set_vec(st, "o_t", k0, o0)
    x = 0.05
    nsteps = 100
    for k in range(1, nsteps + 1):
        nk = _rot(axis, k * x, n0)
        _set_particle(st, "R", 0, c - 0.5 * rho * nk)
        _set_particle(st, "R", 1, c + 0.5 * rho * nk)
        o.set_state(st)
        o.call("ora_forces")
        st = o.get_state()
        ok = _vec(st, "o_t", k0)
        assert abs(np.dot(ok, nk)) < 1e-13 * osz
        assert np.linalg.norm(ok) == pytest.approx(osz, rel=1e-13)
    nN = _rot(axis, nsteps * x, n0)
    ck = math.cos(x) ** nsteps
    exp = osz * (a * axis + b * ck * np.cross(axis, nN)) / math.sqrt(a * a + b * b * ck * ck)
    assert _rel(_vec(st, "o_t", k0), exp) < 1e-11
    Ft = Kt * exp
    assert _rel(_particle(st, "F", 0), nN * fn + Ft) < 1e-11

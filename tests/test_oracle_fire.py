"""Pins of the FIRE oracle (oracle/fire_oracle.c) by closed forms and
invariants, and the C-ABI defaults.

The reference holds no FIRE fixtures; each test below fixes one part of the
restatement by what the mathematics or the reference's constants determine:
the A8 increment (md_firerelaxation.cu:28-52), the dt/alpha schedule
(md_createcylindersample.cu:374-397, SetFIREData md_firerelaxation.cu:72-86),
the cylinder projection (md_cylinderrestriction.cu:84-108), particle removal
(md_reduceparticlenumber.cu:18-39) and convergence to an overlap-free state.
"""
import ctypes as C

import numpy as np
import pytest

import pyoracle
from rockfit import abi

PI = 3.1415926535897932384626433832795
RHO = 2.65  # reduced density (2650 kg/m^3)


def mass(r):
    return RHO * (4.0 / 3.0) * PI * r * r * r


def pair(dist, r=0.25):
    xyz = np.array([[0.0, 0.0, 0.0], [dist, 0.0, 0.0]])
    return xyz, np.array([r, r])


def cfg(**kw):
    c = pyoracle.fire_defaults()
    c.D, c.dt0, c.density = 1.0, 1e-3, RHO
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def test_defaults_are_the_reference_constants():
    c = pyoracle.fire_defaults()
    # SetFIREData (md_firerelaxation.cu:76-84) and alpha0 of the loop (:349)
    assert (c.dtmax_factor, c.dtmin_factor, c.dtgrow, c.dtshrink) == (100.0, 0.001, 1.1, 0.5)
    assert (c.alpha0, c.alpha_shrink, c.ndelay, c.np_negative_max) == (0.1, 0.99, 20, 30000)
    assert c.quirks == abi.RF_QUIRK_FIRE_NO_INERTIA


def test_abi_defaults_match_oracle():
    from rockfit import _lib
    g = abi.FireCfg()
    assert _lib.lib().rf_fire_defaults(C.byref(g)) == 0
    o = pyoracle.fire_defaults()
    assert abi.struct_to_dict(g) == abi.struct_to_dict(o)


def test_no_inertia_increment_is_force_times_dt0_over_m():
    """A8: one step from rest moves each particle by (F dt0/m) dt0; the alpha
    mixing of a velocity parallel to F is the identity (to rounding)."""
    d, r, D, dt0 = 0.4, 0.25, 1.0, 1e-3
    xyz, rr = pair(d, r)
    x1, v1, out = pyoracle.fire_relax(xyz, rr, cfg(max_steps=1))
    assert out.steps == 1 and out.stop_reason == 1
    a = 2 * r
    F0 = -D * (a * a / (d * d)) * d        # force on particle 0 along x
    v = F0 * dt0 / mass(r)
    assert v1[0, 0] == pytest.approx(v, rel=1e-13)
    assert x1[0, 0] == pytest.approx(v * dt0, rel=1e-12)
    assert x1[1, 0] == pytest.approx(d - v * dt0, rel=1e-12)
    assert np.all(x1[:, 1:] == 0.0)
    # first step: F.V = 0 is not positive, so it counts as a negative step
    assert (out.np_positive, out.np_negative) == (0, 1)


def test_dt_and_alpha_schedule():
    """Power stays positive while the pair separates: after the first step,
    every step is positive, and dt grows by 1.1 (capped at 100 dt0) and alpha
    shrinks by 0.99 once more than 20 positive steps have been seen."""
    K, dt0 = 60, 1e-3
    xyz, rr = pair(0.4)
    _, _, out = pyoracle.fire_relax(xyz, rr, cfg(max_steps=K, dt0=dt0))
    assert (out.np_positive, out.np_negative) == (K - 1, 0)
    dt, alpha = dt0, 0.1
    for _ in range(K - 1 - 20):
        dt = dt * 1.1 if dt * 1.1 < 100 * dt0 else 100 * dt0
        alpha *= 0.99
    assert out.dt == dt and out.alpha == alpha


def test_negative_power_shrinks_dt_after_ndelay():
    """Standard FIRE on a stiff pair: when the power turns negative after the
    20th step, dt halves and alpha returns to alpha0."""
    xyz, rr = pair(0.45)
    c = cfg(quirks=0, dt0=0.05, D=50.0, max_steps=200)
    _, _, out = pyoracle.fire_relax(xyz, rr, c)
    assert out.np_negative >= 1 or out.np_positive >= 1
    # dt is dt0 times a power of 1.1 and 0.5 bounded by [1e-3, 100] dt0
    assert 1e-3 * 0.05 <= out.dt <= 100 * 0.05


def test_pair_separates_to_contact_and_keeps_its_centre():
    xyz, rr = pair(0.35)
    x1, _, out = pyoracle.fire_relax(xyz, rr, cfg(quirks=0, dt0=0.02, max_steps=400,
                                                   gap_tol=-1e-6, check_every=1))
    dist = np.linalg.norm(x1[1] - x1[0])
    assert dist >= 0.5 * (1 - 1e-6)
    assert abs(x1[:, 0].mean() - 0.175) < 1e-12
    assert out.min_gap == 0.0


def test_cluster_relaxes_to_no_overlap():
    rng = np.random.default_rng(3)
    n = 300
    rr = rng.uniform(0.2, 0.3, n)
    ang = rng.uniform(0, 2 * np.pi, n)
    rad = 2.0 * np.sqrt(rng.uniform(0, 1, n))
    xyz = np.stack([rad * np.cos(ang), rad * np.sin(ang), rng.uniform(-2, 2, n)], 1)
    x1, _, out = pyoracle.fire_relax(xyz, rr, cfg(quirks=0, dt0=0.02, max_steps=300,
                                                   gap_tol=-0.01, check_every=10), threads=4)
    assert out.stop_reason == 1
    # brute force: no pair overlaps by more than 1 %
    d = np.linalg.norm(x1[:, None] - x1[None], axis=2) + np.eye(n) * 1e9
    gap = (d - (rr[:, None] + rr[None])) / (rr[:, None] + rr[None])
    assert gap.min() >= -0.01


def test_cylinder_projection():
    """A centre outside the cylinder is moved back onto its surface along the
    radius and onto the end plane; z beyond hide_z is left alone."""
    xyz = np.array([[3.0, 4.0, 2.5], [0.0, 0.0, 50.0]])
    rr = np.array([0.25, 0.25])
    c = cfg(max_steps=1, container=abi.RF_CONTAINER_CYLINDER_Z, radius=2.0, half_height=2.0,
            hide_z=10.0)
    x1, _, _ = pyoracle.fire_relax(xyz, rr, c)
    assert x1[0, 0] == pytest.approx(1.2, rel=1e-15)
    assert x1[0, 1] == pytest.approx(1.6, rel=1e-15)
    assert x1[0, 2] == 2.0
    assert np.array_equal(x1[1], [0.0, 0.0, 50.0])


def test_removal_of_the_worst_pair():
    """Two nearly coincident spheres and one far away: once checks may remove
    particles, the higher-index member of the worst pair goes to the hidden
    point (and stays there)."""
    xyz = np.array([[0.0, 0.0, 0.0], [0.01, 0.0, 0.0], [5.0, 0.0, 0.0]])
    rr = np.array([0.25, 0.25, 0.25])
    c = cfg(max_steps=4, gap_tol=-0.5, check_every=1, delete_after=0,
            hidden=(C.c_double * 3)(100.0, 100.0, 100.0))
    x1, v1, out = pyoracle.fire_relax(xyz, rr, c)
    assert out.removed == 1
    assert np.array_equal(x1[1], [100.0, 100.0, 100.0]) and np.all(v1[1] == 0.0)
    assert np.array_equal(x1[2], [5.0, 0.0, 0.0])
    assert out.min_gap == 0.0 and out.stop_reason == 1


def test_harmonic_law_first_step():
    """Harmonic law: force D (a - rho) along the normal.  One inertial step
    from rest moves each sphere by F dt0^2 / m (alpha mixing of a velocity
    parallel to F is the identity), against the reference law's D a^2/rho."""
    d, r, D, dt0 = 0.4, 0.25, 3.0, 1e-3
    xyz, rr = pair(d, r)
    c = cfg(max_steps=1, quirks=0, D=D, dt0=dt0, law=abi.RF_FIRE_LAW_HARMONIC)
    x1, v1, out = pyoracle.fire_relax(xyz, rr, c)
    F0 = -D * (2 * r - d)
    v = F0 * dt0 / mass(r)
    assert v1[0, 0] == pytest.approx(v, rel=1e-13)
    assert x1[0, 0] == pytest.approx(v * dt0, rel=1e-12)
    assert x1[1, 0] == pytest.approx(d - v * dt0, rel=1e-12)
    x2, _, _ = pyoracle.fire_relax(xyz, rr, cfg(max_steps=1, quirks=0, D=D, dt0=dt0))
    ratio = x2[0, 0] / x1[0, 0]
    assert ratio == pytest.approx((2 * r) ** 2 / d / (2 * r - d), rel=1e-12)


def test_harmonic_force_vanishes_at_contact():
    """A pair exactly at contact distance feels no force under either law
    (rho^2 < a^2 fails), and a pair just inside contact under the harmonic
    law feels a force proportional to the overlap."""
    for law in (abi.RF_FIRE_LAW_REFERENCE, abi.RF_FIRE_LAW_HARMONIC):
        xyz, rr = pair(0.5)
        x1, v1, _ = pyoracle.fire_relax(xyz, rr, cfg(max_steps=1, quirks=0, law=law))
        assert np.array_equal(x1, xyz) and np.all(v1 == 0.0)
    dx = []
    for ov in (1e-3, 2e-3):
        xyz, rr = pair(0.5 - ov)
        x1, _, _ = pyoracle.fire_relax(xyz, rr, cfg(max_steps=1, quirks=0,
                                                     law=abi.RF_FIRE_LAW_HARMONIC))
        dx.append(-x1[0, 0])
    assert dx[1] / dx[0] == pytest.approx(2.0, rel=1e-9)


def test_absorbing_wall_removes_only_the_outward_velocity():
    """A sphere pushed through the cylinder wall by its neighbour is put back
    on the wall; with the absorbing container its velocity keeps the
    tangential part and loses the outward normal part, with the reflecting
    container (d_CylinderRestrictionZ) the whole radial-plane velocity is
    reversed."""
    # neighbour at the axis side pushes sphere 1 outward along +x, with a
    # y component that is tangential at the wall point
    xyz = np.array([[1.62, -0.3, 0.0], [1.9, 0.0, 0.0]])
    rr = np.array([0.25, 0.25])
    base = dict(max_steps=1, quirks=0, D=5.0, dt0=0.2, radius=1.9, half_height=5.0, hide_z=50.0)
    xa, va, _ = pyoracle.fire_relax(xyz, rr, cfg(container=abi.RF_CONTAINER_CYLINDER_Z_ABSORB,
                                                 **base))
    xr, vr, _ = pyoracle.fire_relax(xyz, rr, cfg(container=abi.RF_CONTAINER_CYLINDER_Z, **base))
    xn, vn, _ = pyoracle.fire_relax(xyz, rr, cfg(container=abi.RF_CONTAINER_NONE, **base))
    # free motion leaves the cylinder
    assert np.hypot(xn[1, 0], xn[1, 1]) > 1.9
    # both containers put the centre back on the wall at the same point
    assert np.hypot(xa[1, 0], xa[1, 1]) == pytest.approx(1.9, rel=1e-15)
    np.testing.assert_array_equal(xa[1], xr[1])
    u = xa[1, :2] / np.hypot(xa[1, 0], xa[1, 1])
    v = vn[1, :2]
    vrad = v @ u
    assert vrad > 0.0
    np.testing.assert_allclose(va[1, :2], v - vrad * u, rtol=1e-13, atol=1e-16)
    assert abs(va[1, :2] @ u) < 1e-14 * np.abs(v).max()
    np.testing.assert_array_equal(vr[1, :2], -v)
    # the end planes: outward z velocity is removed, inward kept
    xyz = np.array([[0.0, 0.0, 4.72], [0.0, 0.0, 4.98]])
    base.update(radius=5.0, half_height=4.98)
    xa, va, _ = pyoracle.fire_relax(xyz, rr, cfg(container=abi.RF_CONTAINER_CYLINDER_Z_ABSORB,
                                                 **base))
    xn, vn, _ = pyoracle.fire_relax(xyz, rr, cfg(container=abi.RF_CONTAINER_NONE, **base))
    assert xn[1, 2] > 4.98 and xa[1, 2] == 4.98
    assert va[1, 2] == 0.0 and vn[1, 2] > 0.0
    assert va[0, 2] == vn[0, 2] < 0.0


def test_dense_growth_converges_with_harmonic_law():
    """Radius growth of a random packing in a cylinder (the command line's
    sample creation, 0.6 -> 0.92 of the final radius): with the harmonic law
    and absorbing walls every stage converges to overlaps below 0.1 %
    (checked by brute force); the reference law with reflecting walls stalls
    at the minimum step."""
    rng = np.random.default_rng(11)
    R0, H0, rf = 3.0, 4.0, 0.5
    n = int(0.6 * np.pi * (R0 + rf) ** 2 * (2 * H0 + 2 * rf) / (4.0 / 3.0 * np.pi * rf ** 3))
    pts = []
    while len(pts) < n:
        X, Y = rng.uniform(-R0, R0, 2)
        if X * X + Y * Y <= R0 * R0:
            pts.append((X, Y, rng.uniform(-H0, H0)))
    start = np.array(pts)

    def grow(law, container):
        xyz = start.copy()
        for sc in (0.6, 0.75, 0.85, 0.92):
            c = cfg(quirks=0, D=1.0, dt0=0.02, density=2.61, law=law, container=container,
                    radius=R0, half_height=H0, hide_z=1e30, max_steps=50, hard_cap=3000,
                    gap_tol=-1e-3, check_every=20)
            xyz, _, out = pyoracle.fire_relax(xyz, np.full(n, sc * rf), c, threads=4)
            if out.stop_reason != 1:
                return xyz, out, sc
        return xyz, out, None

    xyz, out, failed = grow(abi.RF_FIRE_LAW_HARMONIC, abi.RF_CONTAINER_CYLINDER_Z_ABSORB)
    assert failed is None
    a = 2 * 0.92 * rf
    d = np.linalg.norm(xyz[:, None] - xyz[None], axis=2) + np.eye(n) * 1e9
    assert ((d - a) / a).min() >= -1e-3
    assert np.all(np.hypot(xyz[:, 0], xyz[:, 1]) <= R0 * (1 + 1e-15))
    assert np.all(np.abs(xyz[:, 2]) <= H0)
    _, out_ref, failed_ref = grow(abi.RF_FIRE_LAW_REFERENCE, abi.RF_CONTAINER_CYLINDER_Z)
    assert failed_ref is not None and out_ref.dt == pytest.approx(0.02 * 1e-3)

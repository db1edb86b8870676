"""GPU FIRE relaxation (rf_fire_relax) against the oracle, bit for bit.

The order contract (ascending-partner force sums, fixed power reduction,
worst-pair ties to the lower index) makes every position, velocity and
controller value identical.  Cases: the reference's no-inertia increment
(A8) and standard FIRE, a cylinder container, particle removal, and a tiny
skin so neighbour lists are rebuilt many times.
"""
import ctypes as C

import numpy as np
import pytest

import pyoracle
from conftest import gpu_available
from rockfit import abi, fire

pytestmark = pytest.mark.gpu

RESULT_FIELDS = ("steps", "np_positive", "np_negative", "removed", "dt", "alpha", "min_gap",
                 "power", "stop_reason")


def packing(n, seed, R0=4.0, H0=4.0):
    rng = np.random.default_rng(seed)
    r = rng.uniform(0.2, 0.3, n)
    ang = rng.uniform(0, 2 * np.pi, n)
    rad = R0 * np.sqrt(rng.uniform(0, 1, n))
    xyz = np.stack([rad * np.cos(ang), rad * np.sin(ang), rng.uniform(-H0, H0, n)], 1)
    return xyz, r


def run_both(xyz, r, **kw):
    return run_both_x(xyz, r, **kw)[0]


def run_both_x(xyz, r, **kw):
    g = fire.defaults(**dict(dict(D=1.0, density=2.65), **kw))
    o = pyoracle.fire_defaults()
    for k, v in abi.struct_to_dict(g).items():
        if k in ("center", "hidden"):
            setattr(o, k, (C.c_double * 3)(*v))
        else:
            setattr(o, k, v)
    print(f"case n={len(r)} {kw}", flush=True)
    gx, gv, gout = fire.relax(xyz, r, g)
    ox, ov, oout = pyoracle.fire_relax(xyz, r, o, threads=8)
    for f in RESULT_FIELDS:
        assert getattr(gout, f) == getattr(oout, f), (f, getattr(gout, f), getattr(oout, f))
    assert np.array_equal(gx, ox), float(np.abs(gx - ox).max())
    assert np.array_equal(gv, ov), float(np.abs(gv - ov).max())
    return gout, gx


@pytest.fixture(autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.skip("no GPU")


def test_fire_no_inertia_cylinder():
    xyz, r = packing(4000, 1)
    out = run_both(xyz, r, dt0=0.05, max_steps=500, check_every=25,
                   container=abi.RF_CONTAINER_CYLINDER_Z, radius=4.0, half_height=4.0,
                   hide_z=1e9)
    assert out.steps == 500


def test_fire_standard_free():
    xyz, r = packing(4000, 2)
    out = run_both(xyz, r, dt0=0.02, quirks=0, max_steps=400, gap_tol=-0.01, check_every=10)
    assert out.steps >= 400


def test_fire_removal():
    xyz, r = packing(3000, 3, R0=2.5, H0=2.5)
    out = run_both(xyz, r, dt0=0.05, max_steps=300, gap_tol=-0.3, check_every=20,
                   delete_after=50, hard_cap=300, hidden=(500.0, 500.0, 500.0))
    assert out.removed > 0


def test_fire_tiny_skin_rebuilds():
    xyz, r = packing(2000, 4)
    run_both(xyz, r, dt0=0.02, quirks=0, max_steps=300, check_every=7, skin=0.002)


def test_fire_harmonic_absorbing_growth():
    """The command line's packing growth (harmonic law, absorbing walls) at
    two radii and the final stage with removal: bitwise equal at each stage."""
    rng = np.random.default_rng(12)
    R0, H0, n = 6.0, 8.0, 2500  # packing fraction 0.58 at the final radius
    ang = rng.uniform(0, 2 * np.pi, n)
    rad = R0 * np.sqrt(rng.uniform(0, 1, n))
    xyz = np.stack([rad * np.cos(ang), rad * np.sin(ang), rng.uniform(-H0, H0, n)], 1)
    for sc, last in ((0.6, False), (0.85, False), (1.0, True)):
        out, xyz = run_both_x(xyz, np.full(n, 0.5 * sc), dt0=0.02, quirks=0, density=2.61,
                              law=abi.RF_FIRE_LAW_HARMONIC,
                              container=abi.RF_CONTAINER_CYLINDER_Z_ABSORB, radius=R0,
                              half_height=H0, hide_z=1e9, max_steps=50,
                              hard_cap=1500 if last else 3000,
                              gap_tol=-1e-4 if last else -1e-3, check_every=50 if last else 20,
                              delete_after=500 if last else -1, hidden=(500.0, 500.0, 500.0))
        assert out.steps > 0

"""rf_fire_relax_noalloc (no allocation: working memory in the library's
static device array) against rf_fire_relax through the C-ABI.

Bar: positions, velocities and every field of the result bit for bit, on
two calls in a row, the second with fewer particles (the array keeps the
first call's contents); a run larger than the array is refused with
RF_ERR_ARG before any device work.
"""
import numpy as np
import pytest

from rockfit import abi, fire
from rockfit._lib import RockfitError
from test_gpu_fire import RESULT_FIELDS, packing

pytestmark = pytest.mark.gpu


def _cfg():
    return fire.defaults(D=1.0, dt0=0.005, density=2.65, quirks=0, max_steps=300, hard_cap=1500,
                         container=abi.RF_CONTAINER_CYLINDER_Z_ABSORB, center=(0.0, 0.0, 0.0),
                         radius=4.0, half_height=4.0, hide_z=1e30, hidden=(0.0, 0.0, 1e6),
                         gap_tol=-1e-3, check_every=20, delete_after=200,
                         law=abi.RF_FIRE_LAW_HARMONIC)


def test_noalloc_equals_allocating_entry_point():
    for n, seed in ((3000, 21), (1100, 22)):
        print(f"case noalloc n={n} seed={seed}", flush=True)
        xyz, r = packing(n, seed)
        v0 = np.random.default_rng(seed).normal(0.0, 1e-3, (n, 3))
        ax, av, aout = fire.relax(xyz, r, _cfg(), vel=v0)
        bx, bv, bout = fire.relax(xyz, r, _cfg(), vel=v0, noalloc=True)
        for f in RESULT_FIELDS:
            assert getattr(aout, f) == getattr(bout, f), (n, f, getattr(aout, f), getattr(bout, f))
        assert np.array_equal(ax, bx), (n, float(np.abs(ax - bx).max()))
        assert np.array_equal(av, bv), (n, float(np.abs(av - bv).max()))
        assert aout.steps > 20


def test_noalloc_rejects_runs_beyond_the_array():
    xyz, r = packing(200000, 23, R0=20.0, H0=20.0)
    with pytest.raises(RockfitError) as e:
        fire.relax(xyz, r, _cfg(), noalloc=True)
    assert e.value.status == abi.RF_ERR_ARG
    assert "static array" in str(e.value)

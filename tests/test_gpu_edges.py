"""Edge cases of the GPU path, against the CPU oracle.

- Tiny and ragged samples: 1 sphere (no partner, platens only), 2, 33 and
  65 spheres (a last block holding one particle), the spheres nearest the
  centre of a relaxed packing: lists, bonds, state and controller bitwise
  equal over 300 loading steps.
- An interaction row that overflows its slots (8 slots in a dense
  packing): both sides refuse the build with RF_ERR_OVERFLOW
  (md_interaction_list_construct.cu:17-130 writes past the row instead).
- A particle outside the cell grid: both sides refuse the build with
  RF_ERR_GRID (the reference reads outside its cell arrays there).
- Bad arguments (no particles, a radius that is not positive): RF_ERR_ARG.
- The simple-cubic lattice of tests/test_oracle_macro.py (equal distances
  everywhere, so the ranking rules meet exact ties) loaded by the platens:
  bitwise equal to the oracle over 5000 steps.
"""
import numpy as np
import pytest

import pyoracle
from helpers import compare_states, compare_test_states
from rockfit import abi, config
from rockfit._lib import RockfitError
from rockfit.core import Context

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def packing():
    return config.make_workload(n=600, r_min=0.2, r_max=0.3, test="uniaxial")


def _subset(w, n_keep):
    c = w["xyz"].mean(axis=0)
    idx = np.argsort(np.linalg.norm(w["xyz"] - c, axis=1), kind="stable")[:n_keep]
    xyz, r = w["xyz"][idx], w["r"][idx]
    xyz, box = config.place_in_box(xyz, r, w["mat"])
    cfg = config.uniaxial_cfg(xyz, r, v_load_mps=-2.0, dnsave=5, max_strain=0.05)
    return dict(xyz=xyz, r=r, mat=w["mat"], box=box, cfg=cfg)


@pytest.mark.parametrize("n", [1, 2, 33, 65])
def test_tiny_and_ragged_samples_parity(packing, n):
    w = _subset(packing, n)
    print(f"case tiny sample n={n}", flush=True)
    g = Context(w["xyz"], w["r"], w["mat"], w["box"])
    o = pyoracle.Oracle(w["xyz"], w["r"], w["mat"], w["box"], threads=1)
    g.build_bonds()
    o.build_bonds()
    compare_states(g.get_state(), o.get_state(), g.n, g.slots, what=f"n={n} build")
    g.set_test(w["cfg"])
    o.set_test(w["cfg"])
    done = 0
    for chunk in (1, 99, 200):
        kg, ko = g.step(chunk), o.step(chunk)
        assert kg == ko
        done += kg
        compare_states(g.get_state(), o.get_state(), g.n, g.slots, what=f"n={n} step {done}")
        compare_test_states(g.test_state(), o.test_state(), what=f"n={n} step {done}")
    np.testing.assert_array_equal(g.curve(), o.curve())
    np.testing.assert_array_equal(g.bond_counts(), o.bond_counts())


def _oracle_status(fn):
    with pytest.raises(RuntimeError) as e:
        fn()
    return str(e.value)


def test_row_overflow_is_refused(packing):
    w = packing
    print("case row overflow slots=8", flush=True)
    g = Context(w["xyz"], w["r"], w["mat"], w["box"], slots=8)
    with pytest.raises(RockfitError) as e:
        g.build_bonds()
    assert e.value.status == abi.RF_ERR_OVERFLOW
    o = pyoracle.Oracle(w["xyz"], w["r"], w["mat"], w["box"], slots=8, threads=1)
    assert f"status {abi.RF_ERR_OVERFLOW}" in _oracle_status(o.build_bonds)


def test_particle_outside_the_grid_is_refused(packing):
    w = packing
    print("case particle outside the cell grid", flush=True)
    xyz = w["xyz"].copy()
    xyz[7, 0] = -3.0 * w["box"].cell  # three cells below the grid's first
    g = Context(xyz, w["r"], w["mat"], w["box"])
    with pytest.raises(RockfitError) as e:
        g.build_bonds()
    assert e.value.status == abi.RF_ERR_GRID
    o = pyoracle.Oracle(xyz, w["r"], w["mat"], w["box"], threads=1)
    assert f"status {abi.RF_ERR_GRID}" in _oracle_status(o.build_bonds)


def test_bad_arguments_are_refused(packing):
    w = packing
    print("case bad arguments", flush=True)
    with pytest.raises(RockfitError) as e:
        Context(w["xyz"][:0], w["r"][:0], w["mat"], w["box"])
    assert e.value.status == abi.RF_ERR_ARG
    r = w["r"].copy()
    r[3] = 0.0
    with pytest.raises(RockfitError) as e:
        Context(w["xyz"], r, w["mat"], w["box"])
    assert e.value.status == abi.RF_ERR_ARG


def test_simple_cubic_lattice_parity():
    import test_oracle_macro as lattice

    xyz, r, mat, box, cfg = lattice._lattice()
    print("case simple-cubic lattice", flush=True)
    g = Context(xyz, r, mat, box)
    o = pyoracle.Oracle(xyz, r, mat, box, threads=1)
    g.build_bonds()
    o.build_bonds()
    compare_states(g.get_state(), o.get_state(), g.n, g.slots, what="lattice build")
    g.set_test(cfg)
    o.set_test(cfg)
    for chunk in (1000, 4000):
        assert g.step(chunk) == o.step(chunk) == chunk
    compare_states(g.get_state(), o.get_state(), g.n, g.slots, what="lattice 5000 steps")
    compare_test_states(g.test_state(), o.test_state(), what="lattice 5000 steps")
    np.testing.assert_array_equal(g.curve(), o.curve())

"""GPU step against the oracle with every history term live, and the graph
cache across a second rf_set_test.

The loading workloads start from zero bond history with prev_n = 0
(md_boundinteractions_construct.cu:69-80), so their first steps leave most
of M(t, phi) (md_calculateforcesdem3.cu:57-101) multiplying zeros.  Here
every bond slot gets a previous normal turned ~0.2 rad away from the
current one, random shear, torsion and bending histories, and every
particle a spin of ~0.01 rad per step (bonds are unbreakable here, so
the histories stay live for every step), so each of the nine entries of M (and
its second-order part) multiplies non-zero history on both sides; contact
slots get random stored overlaps, not in the tangent plane, so the
projection and rescale (:264-279) act.
"""
import numpy as np
import pytest

import pyoracle
from helpers import compare_states, compare_test_states
from rockfit import config
from rockfit.core import Context

pytestmark = pytest.mark.gpu


def _perturbed_state(st, n, S, seed):
    rng = np.random.default_rng(seed)
    R = st["R"].reshape(3, n).T
    nbr = st["nbr"]
    occ = np.nonzero(nbr >= 0)[0]
    i = occ // S
    j = nbr[occ]
    rv = R[j] - R[i]
    nv = rv / np.linalg.norm(rv, axis=1)[:, None]
    perp = np.cross(nv, rng.normal(size=nv.shape))
    perp /= np.linalg.norm(perp, axis=1)[:, None]
    pn = nv + 0.2 * perp
    pn /= np.linalg.norm(pn, axis=1)[:, None]
    bond = st["type"][occ] == 1
    for f, scale in (("prev_n", None), ("o_t", 2e-5), ("m_n", 2e-4), ("m_t", 2e-4)):
        a = st[f].reshape(-1, 3)
        if f == "prev_n":
            a[occ[bond]] = pn[bond]
        else:
            a[occ[bond]] = rng.normal(0.0, scale, (int(bond.sum()), 3))
    # contact slots: stored overlap with a part along the normal
    a = st["o_t"].reshape(-1, 3)
    cont = ~bond
    a[occ[cont]] = rng.normal(0.0, 1e-5, (int(cont.sum()), 3))
    return st


def test_random_history_parity():
    w = config.make_workload(n=1500, r_min=0.2, r_max=0.3, test="uniaxial", seed=77,
                             mat_kwargs=dict(sigma_mpa=1e9, tau_mpa=1e9),
                             test_kwargs=dict(v_load_mps=-1.0, dnsave=2, max_strain=0.05))
    print("case random history n=1500 seed=77", flush=True)
    g = Context(w["xyz"], w["r"], w["mat"], w["box"])
    o = pyoracle.Oracle(w["xyz"], w["r"], w["mat"], w["box"], threads=8)
    g.build_bonds()
    o.build_bonds()
    d = pyoracle.derive(w["mat"], float(w["r"].min()), float(w["r"].max()))
    st = _perturbed_state(o.get_state(), o.n, o.S, 5)
    rng = np.random.default_rng(6)
    st["W"][:] = rng.normal(0.0, 0.01 / d.dt, 3 * o.n)
    g.set_state(st)
    o.set_state(st)
    g.set_test(w["cfg"])
    o.set_test(w["cfg"])
    done = 0
    nb0 = o.counters().bonds
    for chunk in (1, 1, 3):
        assert g.step(chunk) == o.step(chunk) == chunk
        done += chunk
        compare_states(g.get_state(), o.get_state(), g.n, g.slots, what=f"random history +{done}")
        compare_test_states(g.test_state(), o.test_state(), what=f"random history +{done}")
    assert o.counters().bonds == nb0 > 4 * o.n


def test_set_test_twice_replays_fresh_buffers():
    """set_test -> graphed chunks -> set_state -> set_test (same cfg) ->
    graphed chunks equals a fresh context: the second set_test reallocates
    the curve and bond-count buffers, which captured graphs must not keep
    writing to."""
    w = config.make_workload(n=1500, r_min=0.2, r_max=0.3, test="uniaxial",
                             test_kwargs=dict(v_load_mps=-2.0, dnsave=5, max_strain=0.05))
    print("case set_test twice n=1500", flush=True)
    g = Context(w["xyz"], w["r"], w["mat"], w["box"])
    g.build_bonds()
    g.set_chunk(64)
    st0 = g.get_state()
    g.set_test(w["cfg"])
    assert g.step(256) == 256
    g.set_state(st0)
    g.set_test(w["cfg"])
    assert g.step(256) == 256
    f = Context(w["xyz"], w["r"], w["mat"], w["box"])
    f.build_bonds()
    f.set_chunk(64)
    f.set_test(w["cfg"])
    assert f.step(256) == 256
    np.testing.assert_array_equal(g.curve(), f.curve())
    np.testing.assert_array_equal(g.bond_counts(), f.bond_counts())
    compare_test_states(g.test_state(), f.test_state(), what="second set_test")
    compare_states(g.get_state(), f.get_state(), g.n, g.slots, what="second set_test")
    assert np.any(g.curve()[:, 2] != 0.0)

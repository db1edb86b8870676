"""Pin the oracle's cell sort and interaction-list maintenance.

* cell order = numpy's stable argsort of the cell keys (any stable sort by
  key equals cub::DeviceRadixSort::SortPairs over identity values);
* rows after construction / append equal the brute-force O(N^2) pair set
  {j : 1e-10 < |r_j - r_i|^2 < acut^2};
* the reference's slot rules: contiguous construction order, first free slot
  on append, particle 0's sentinel quirk, overflow reported as an error.
"""
import numpy as np
import pytest

import pyoracle
from helpers import randomize_motion
from rockfit import abi, config


def _brute_pairs(xyz, acut):
    d = xyz[:, None, :] - xyz[None, :, :]
    rr = np.einsum("ijk,ijk->ij", d, d)
    ok = (rr < acut * acut) & (rr > 1e-10)
    return {(i, j) for i, j in zip(*np.nonzero(ok))}


def _row_pairs(st, n, S):
    nbr = st["nbr"].reshape(n, S)
    return {(i, int(j)) for i in range(n) for j in nbr[i] if j >= 0}


def _xyz(st, n):
    return st["R"].reshape(3, n).T.copy()


def working_grid(b):
    """The reference grid (dims) plus one cell layer past the top x and y
    faces, which the oracle and the kernels index with."""
    return b.dims[0] + 1, b.dims[1] + 1


@pytest.fixture(scope="module")
def wl():
    return config.make_workload(n=600, r_min=0.2, r_max=0.3, test="none",
                                gen_kwargs=dict(max_iter=100))


def test_cell_sort_is_stable_by_key(wl):
    o = pyoracle.Oracle(wl["xyz"], wl["r"], wl["mat"], wl["box"])
    key, srt, st, ct = o.cells()
    order = np.argsort(key, kind="stable")
    assert np.array_equal(srt, order)
    assert np.array_equal(np.bincount(key, minlength=len(ct)), ct)
    assert np.array_equal(np.concatenate([[0], np.cumsum(ct)[:-1]]), st)
    b = wl["box"]
    a = b.cell
    ix = np.floor(wl["xyz"] / a).astype(int)
    nx, ny = working_grid(b)
    assert np.array_equal(key, ix[:, 0] + ix[:, 1] * nx + ix[:, 2] * nx * ny)
    assert len(ct) == nx * ny * b.dims[2]


def test_construct_rows_equal_bruteforce(wl):
    o = pyoracle.Oracle(wl["xyz"], wl["r"], wl["mat"], wl["box"])
    o.build_bonds()
    st = o.get_state()
    n, S = o.n, o.S
    assert _row_pairs(st, n, S) == _brute_pairs(wl["xyz"], wl["mat"].acut)
    # rows are contiguous from slot 0 right after construction
    occ = st["nbr"].reshape(n, S) >= 0
    cnt = occ.sum(axis=1)
    assert np.all(occ == (np.arange(S)[None, :] < cnt[:, None]))


def test_construct_row_order_follows_stencil(wl):
    """Slot order = stencil cell order, ascending id inside a cell."""
    o = pyoracle.Oracle(wl["xyz"], wl["r"], wl["mat"], wl["box"])
    o.build_bonds()
    st = o.get_state()
    key, _, _, _ = o.cells()
    b = wl["box"]
    a = b.cell
    nx, ny = working_grid(b)
    nxy = nx * ny
    xyz = wl["xyz"]
    for i in range(0, o.n, 37):
        f = np.floor(xyz[i] / a)
        s = np.where(np.copysign(1.0, xyz[i] - (f + 0.5) * a) > 0, 1, -1)
        base = int(f[0] + f[1] * nx + f[2] * nxy)
        cells = [base, base + s[0], base + s[1] * nx, base + s[2] * nxy,
                 base + s[0] + s[1] * nx, base + s[0] + s[2] * nxy,
                 base + s[1] * nx + s[2] * nxy, base + s[0] + s[1] * nx + s[2] * nxy]
        rank = {c: p for p, c in enumerate(cells)}
        row = [j for j in st["nbr"].reshape(o.n, o.S)[i] if j >= 0]
        keys = [(rank[int(key[j])], j) for j in row]
        assert keys == sorted(keys)


def test_append_and_prune_track_bruteforce(wl):
    o = pyoracle.Oracle(wl["xyz"], wl["r"], wl["mat"], wl["box"])
    o.build_bonds()
    st = randomize_motion(o.get_state(), o.n, 3, 3e-3, 0.0)
    o.set_state(st)
    cfg = abi.PlatenCfg()
    cfg.kind = abi.RF_TEST_NONE
    o.set_test(cfg)
    for _ in range(4):
        o.step(25)
        o.call("ora_append")  # the next step's append, on the current positions
        s = o.get_state()
        assert _row_pairs(s, o.n, o.S) == _brute_pairs(_xyz(s, o.n), wl["mat"].acut)


def _line_sample(positions, mat, space=(12.0, 12.0, 12.0)):
    box = config.box_reference(list(space), 2.0 * mat.acut)
    c = np.array(box.center[:])
    xyz = c + np.asarray(positions, float)
    r = np.full(len(xyz), 0.25)
    return xyz, r, box


@pytest.mark.parametrize("quirks,expect_slot", [(abi.RF_QUIRKS_REFERENCE, 2), (0, 0)])
def test_first_free_slot_and_particle0_sentinel(quirks, expect_slot):
    """Particle 0 with slot 0 and slot 2 empty receives a new partner in slot
    2 with the reference's sentinel quirk, in slot 0 without it."""
    mat = config.material(0.25, 0.25, gap_factor=-1.0, quirks=quirks)
    pos = [(0, 0, 0), (0.6, 0, 0), (-0.6, 0, 0), (0, 0.6, 0), (0, 3.0, 0)]
    xyz, r, box = _line_sample(pos, mat)
    o = pyoracle.Oracle(xyz, r, mat, box, slots=8)
    o.build_bonds()
    st = o.get_state()
    S = o.S
    row0 = st["nbr"][:S].copy()
    assert sorted(row0[row0 >= 0].tolist()) == [1, 2, 3]
    # move particle 0's partners to slots 1, 3, 4 (slots 0 and 2 empty)
    partners = row0[row0 >= 0].tolist()
    st["nbr"][:S] = -1
    for k, j in zip((1, 3, 4), partners):
        st["nbr"][k] = j
    # bring particle 4 within acut of particle 0: the only new pair
    n = o.n
    st["R"][n + 4] = st["R"][n + 0] + 0.7
    o.set_state(st)
    o.call("ora_prune")
    o.call("ora_append")
    row = o.get_state()["nbr"][:S]
    assert row[expect_slot] == 4
    empty = 0 if expect_slot == 2 else 2
    assert row[empty] == -1
    assert sorted(row[[1, 3, 4]].tolist()) == sorted(partners)


def test_row_overflow_is_an_error():
    mat = config.material(0.25, 0.25, gap_factor=-1.0)
    rng = np.random.default_rng(1)
    pos = rng.uniform(-0.3, 0.3, size=(12, 3))
    xyz, r, box = _line_sample(pos, mat)
    o = pyoracle.Oracle(xyz, r, mat, box, slots=4)
    with pytest.raises(RuntimeError, match="overflow"):
        o.build_bonds()


def _morton_order(xyz, cell):
    s = 4.0 * (1.0 / cell)
    q = np.clip(np.floor(xyz * s), 0, 2097151).astype(np.uint64)
    key = np.zeros(len(xyz), dtype=np.uint64)
    for b in range(21):
        for a in range(3):
            key |= ((q[:, a] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + a)
    return np.lexsort((np.arange(len(xyz)), key))


def test_platen_sum_order():
    """The oracle groups the platen-force sum like the device stores its
    particles: by the Morton code of the quarter-cell coordinates
    floor(4 x / cell), ties by id (RF_ORDER_SPATIAL), or by id
    (RF_ORDER_GIVEN).  Restated here with numpy bit operations."""
    w = config.make_workload(n=1200, r_min=0.2, r_max=0.3, test="uniaxial", seed=4)
    o = pyoracle.Oracle(w["xyz"], w["r"], w["mat"], w["box"])
    np.testing.assert_array_equal(o.order(), _morton_order(w["xyz"], w["box"].cell))
    assert not np.array_equal(o.order(), np.arange(o.n))
    g = pyoracle.Oracle(w["xyz"], w["r"], w["mat"], w["box"], order=abi.RF_ORDER_GIVEN)
    np.testing.assert_array_equal(g.order(), np.arange(g.n))
    # the two orders group the same platen forces differently: equal to rounding
    cfg = config.uniaxial_cfg(w["xyz"], w["r"], v_load_mps=-2.0, dnsave=1)
    for x in (o, g):
        x.build_bonds()
        x.set_test(cfg)
        x.step(150)
    a, b = o.test_state(), g.test_state()
    assert a.sfl1 != 0.0 and abs(a.sfl1 - b.sfl1) <= 1e-12 * abs(a.sfl1)


def _tree(v):
    v = list(v)
    off = len(v) // 2
    while off >= 1:
        for t in range(off):
            v[t] = v[t] + v[t + off]
        off //= 2
    return v[0]


def test_platen_group_tree():
    """The group partials of the platen-force sum: per 32 particles in
    storage order (-0.0 past the last), each run of 8 by the tree
    v[k] += v[k+off], off = 4, 2, 1, then the four runs as
    (W0 + W2) + (W1 + W3) -- the device's waves and controller.  Restated
    in numpy float64 and compared bit for bit, with platens pressed into the
    specimen so that many particles carry distinct forces; a plain 32-wide
    tree gives different bits, so the check tells the orders apart."""
    w = config.make_workload(n=1200, r_min=0.2, r_max=0.3, test="uniaxial", seed=6)
    o = pyoracle.Oracle(w["xyz"], w["r"], w["mat"], w["box"])
    o.build_bonds()
    cfg = w["cfg"]
    cfg.size[2] -= 0.1  # both platens 0.05 mm inside the envelope
    o.set_test(cfg)
    o.call("ora_forces")
    pb, pt = o.particle_partials()
    gb, gt = o.partials()
    order = o.order()
    n = o.n
    assert np.count_nonzero(pt) > 25 and np.count_nonzero(pb) > 25
    differs = False
    for per, got in ((pb, gb), (pt, gt)):
        for g in range(len(got)):
            v = [float(per[order[32 * g + j]]) if 32 * g + j < n else -0.0 for j in range(32)]
            W = [_tree(v[8 * r:8 * r + 8]) for r in range(4)]
            expect = (W[0] + W[2]) + (W[1] + W[3])
            assert np.float64(got[g]).tobytes() == np.float64(expect).tobytes(), g
            differs |= np.float64(_tree(v)).tobytes() != np.float64(expect).tobytes()
    assert differs

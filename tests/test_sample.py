"""Sample generator and the reference's Sample_*.txt format."""
import numpy as np
import pytest
from scipy.spatial import cKDTree

import pyoracle
from rockfit import sample


def test_generator_is_seeded_and_dense():
    a = sample.generate(800, 0.2, 0.3, seed=11)
    b = sample.generate(800, 0.2, 0.3, seed=11)
    c = sample.generate(800, 0.2, 0.3, seed=12)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert not np.array_equal(a[0], c[0])
    xyz, r = a
    assert r.min() >= 0.2 and r.max() <= 0.3
    t = cKDTree(xyz)
    pr = t.query_pairs(0.6, output_type="ndarray")
    gap = np.linalg.norm(xyz[pr[:, 1]] - xyz[pr[:, 0]], axis=1) - r[pr[:, 0]] - r[pr[:, 1]]
    # many near-touching pairs: about 4-7 bonds per particle at gap < 0.05 r_min
    bonds = 2 * np.count_nonzero(gap < 0.05 * 0.2) / len(r)
    assert 3.5 < bonds < 8.0
    assert -gap.min() < 0.2 * 0.2


def test_sample_file_roundtrip(tmp_path):
    xyz, r = sample.generate(100, 0.2, 0.3, seed=5, max_iter=20)
    p = tmp_path / "Sample_0_0.txt"
    sample.write_sample(p, xyz, r)
    x2, r2 = sample.read_sample(p)
    # x * 1e-3 written with 17 digits, read back as x_m * (1 / 1e-3)
    assert np.allclose(x2, xyz, rtol=1e-15, atol=0)
    assert np.allclose(r2, r, rtol=1e-15, atol=0)


def test_reference_three_column_file(tmp_path):
    p = tmp_path / "Sample_0_1.txt"
    p.write_text("3\n0.001 0.002 0.003\n0.0015 0.002 0.003\n-0.001 0 1e-3\n")
    xyz, r = sample.read_sample(p, default_radius_mm=0.5)
    assert xyz.shape == (3, 3)
    assert xyz[0, 0] == pytest.approx(1.0) and xyz[2, 2] == pytest.approx(1.0)
    assert np.all(r == 0.5)
    with pytest.raises(ValueError):
        bad = tmp_path / "bad.txt"
        bad.write_text("2\n0 0 0\n")
        sample.read_sample(bad)


def test_bond_strength_draw_statistics():
    L = pyoracle.lib()
    vals = np.array([L.ora_bond_strength(99, i, i + 7, 0, 40.0, 0.2) for i in range(20000)])
    assert vals.mean() == pytest.approx(40.0, rel=0.01)
    assert vals.std() / vals.mean() == pytest.approx(0.2, rel=0.03)
    # log-normal: log values are normal with variance ln(1 + cv^2)
    assert np.log(vals).std() == pytest.approx(np.sqrt(np.log(1.04)), rel=0.03)
    assert L.ora_bond_strength(99, 3, 10, 0, 40.0, 0.2) == L.ora_bond_strength(99, 3, 10, 0, 40.0, 0.2)
    assert L.ora_bond_strength(99, 3, 10, 0, 40.0, 0.0) == 40.0


def test_random_sequential_insertion():
    """RSI places every sphere inside the container without overlaps, in a
    seeded, repeatable way."""
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(3)
    n = 6000
    r = rng.uniform(0.12, 0.18, n)
    Rc, Hc = 3.0, 6.0
    p = sample._rsi(r, Rc, Hc, np.random.default_rng(9))
    assert np.all(np.hypot(p[:, 0], p[:, 1]) <= Rc) and np.all(np.abs(p[:, 2]) <= Hc)
    pairs = cKDTree(p).query_pairs(2 * r.max(), output_type="ndarray")
    d = np.linalg.norm(p[pairs[:, 0]] - p[pairs[:, 1]], axis=1)
    assert np.all(d >= r[pairs[:, 0]] + r[pairs[:, 1]])
    np.testing.assert_array_equal(p, sample._rsi(r, Rc, Hc, np.random.default_rng(9)))

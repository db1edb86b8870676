"""Pin the oracle's macro response: a simple-cubic bonded lattice loaded by
the uniaxial platens.

A lattice of equal spheres (radius r, spacing d = 2r, 3 x 3 columns of 12)
bonds each sphere to its touching neighbours.  Loaded along the columns,
only the axial bonds carry load. Each column is then n - 1 bonds in series,
each of stiffness k = E_b pi (ratio r)^2 / d (the elastic limit of the
lattice: E_macro = E_b pi R_b^2 / d^2), between two Hertz contacts with the
platens, F = 4/3 E/(1 - nu^2) sqrt(r) delta^1.5 (wall moduli 2 hm_E,
md_calculationuniaxialcompression.cu:278).  At a curve point with bottom
platen stress s, each column carries F = s A / 9, with A = pi (R + r)^2 the
reference's area (md_data_types.h:212-242). The platens have then closed by
dH = (n - 1) F / k + 2 delta(F), which the curve's strain gives as
-strain * size0 (md_results.cu:56-99, md_data_types.h:347-379).

Loading at 0.001 m/s is quasi-static after the first contact transient: the
second half of the curve must satisfy the closed form to 2e-3 relative.
tests/test_oracle_mutants.py checks that a wrong stress area, a platen
moving a full instead of a half step, a doubled bond cross-section or a
halved wall modulus fail this module.
"""
import math

import numpy as np

import pyoracle
from rockfit import config

R = 0.5
NX, NZ = 3, 12
STEPS = 30000


def _lattice():
    g = (np.arange(NX) - (NX - 1) / 2) * 2 * R
    xyz = np.array([(x, y, z * 2 * R) for z in range(NZ) for y in g for x in g], dtype=float)
    r = np.full(len(xyz), R)
    # unbreakable bonds: the pin is the elastic response
    mat = config.material(R, R, sigma_mpa=1e6, tau_mpa=1e6)
    xyz, box = config.place_in_box(xyz, r, mat)
    cfg = config.uniaxial_cfg(xyz, r, v_load_mps=-0.001, dnsave=200, max_strain=0.5,
                              min_stress_mpa=1e4)
    return xyz, r, mat, box, cfg


def test_simple_cubic_lattice_macro_response():
    xyz, r, mat, box, cfg = _lattice()
    o = pyoracle.Oracle(xyz, r, mat, box, threads=1)
    o.build_bonds()
    st = o.get_state()
    # axial bonds 9 x (NZ - 1), lateral 12 per layer, both directions
    assert np.count_nonzero(st["type"] == 1) == 2 * (NX * NX * (NZ - 1) + 2 * NX * (NX - 1) * NZ)
    o.set_test(cfg)
    assert o.step(STEPS) == STEPS
    curve = o.curve()
    area = math.pi * (0.5 * cfg.size[0] + cfg.r_ref) ** 2
    k = mat.E_bond * math.pi * (mat.bond_radius_ratio * R) ** 2 / (2 * R)
    hertz = 4.0 / 3.0 * mat.E_contact / (1.0 - mat.nu ** 2) * math.sqrt(R)
    pts = curve[len(curve) // 2:]
    assert len(pts) >= 50
    for strain, s_b, s_t in pts:
        assert abs(abs(s_b) - abs(s_t)) <= 1e-3 * abs(s_b)  # the column is balanced
        F = abs(s_b) * area / (NX * NX)
        dH = (NZ - 1) * F / k + 2.0 * (F / hertz) ** (2.0 / 3.0)
        assert abs(-strain * cfg.size[2] / dH - 1.0) < 2e-3, (strain, s_b, dH)
    # the last point is well into the loading (about 0.1 MPa)
    assert abs(pts[-1][1]) * 1e3 > 0.05

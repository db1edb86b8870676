"""Derivatives of the material constants that seed a sensitivity run
(rf_sens_material_tangent) against central differences of rf_derive
(potential_data::CalculateParameters, md_data_types.h:155-187)."""
import numpy as np
import pytest

from rockfit import abi, config
from rockfit.core import derive, material_tangent

FIELDS = ("E_bond", "nu", "E_contact", "friction", "rolling_friction", "bond_radius_ratio")
PARAM = dict(zip(FIELDS, range(6)))


def constants(m, wf):
    d = derive(m, 0.2, 0.3)
    return np.array([m.E_bond, d.G_bond, d.hm_E, d.hm_G, m.friction, m.rolling_friction, d.p_A,
                     wf * d.hm_E, wf * d.hm_G, m.bond_radius_ratio, m.bond_radius])


@pytest.mark.parametrize("field", FIELDS)
def test_material_tangent_is_the_derivative_of_rf_derive(field):
    m = config.material(0.2, 0.25, E_bond_gpa=20.0, E_contact_gpa=50.0, nu=0.25)
    wf = 2.0
    got = material_tangent(m, PARAM[field], wf)
    x0 = getattr(m, field)
    h = 1e-6 * abs(x0)
    mp = abi.Material.from_buffer_copy(m)
    mm = abi.Material.from_buffer_copy(m)
    setattr(mp, field, x0 + h)
    setattr(mm, field, x0 - h)
    fd = (constants(mp, wf) - constants(mm, wf)) / (2 * h)
    np.testing.assert_allclose(got, fd, rtol=1e-7, atol=1e-9 * np.abs(fd).max())

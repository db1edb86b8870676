"""The C-ABI library: loads without a GPU, exports every declared symbol,
struct layouts agree between C and the ctypes mirrors, and the host-only
functions (rf_derive, rf_box_reference) equal the oracle's bitwise."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import pyoracle
from rockfit import _lib, abi, config, core

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rockfit_dem.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rf_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert "rf_create" in names and "rf_step" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, f"not exported: {missing}"
    assert set(_lib.EXPORTS) <= set(names)
    assert L.rf_version() == 1


@pytest.mark.parametrize("which,cls", [(0, abi.Material), (1, abi.Box), (2, abi.Sample),
                                       (3, abi.Options), (4, abi.Derived), (5, abi.PlatenCfg),
                                       (6, abi.TestState), (7, abi.Counters),
                                       (8, abi.StateView), (9, abi.FireCfg),
                                       (10, abi.FireResult), (11, abi.Result),
                                       (12, abi.AdjointCfg)])
def test_struct_layout(which, cls):
    assert pyoracle.lib().ora_sizeof(which) == C.sizeof(cls)


def test_derive_matches_oracle_and_closed_forms():
    m = config.material(0.2, 0.25)
    d = core.derive(m, 0.2, 0.3)
    o = pyoracle.derive(m, 0.2, 0.3)
    assert abi.struct_to_dict(d) == abi.struct_to_dict(o)
    G = m.E_contact / (2 * (1 + m.nu))
    # Rayleigh time pi r sqrt(rho/G) / (0.163 nu + 0.8766)
    tr = np.pi * 0.2 * np.sqrt(m.density / G) / (0.163 * m.nu + 0.8766)
    assert d.T_rayleigh == pytest.approx(tr, rel=1e-14)
    assert d.hm_E == pytest.approx(m.E_contact / (2 * (1 - m.nu ** 2)), rel=1e-15)
    assert d.p_A == pytest.approx(np.log(0.6) / np.hypot(np.pi, np.log(0.6)), rel=1e-15)
    assert d.dt == pytest.approx(0.05 * min(d.T_rayleigh, d.T_bond), rel=1e-15)


def test_box_reference_matches():
    for space, cell in (([40.0, 40.0, 60.0], 3.2), ([13.1, 9.7, 22.3], 1.6)):
        a = core.box_reference(space, cell)
        b = pyoracle.box_reference(space, cell)
        c = config.box_reference(space, cell)
        assert abi.struct_to_dict(a) == abi.struct_to_dict(b) == abi.struct_to_dict(c)
        # the hidden point lies inside the padded grid, above the space box
        iz = int(np.floor(a.hidden[2] / cell))
        assert 1 <= iz <= a.dims[2] - 2
        assert a.hidden[2] > a.center[2] + a.half[2]


def test_create_without_device_fails_loudly():
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    w = config.make_workload(n=50, r_min=0.2, r_max=0.3, test="none",
                             gen_kwargs=dict(max_iter=10))
    with pytest.raises(_lib.RockfitError, match="RF_ERR_NODEVICE"):
        core.Context(w["xyz"], w["r"], w["mat"], w["box"])
    from rockfit import fire
    c = fire.defaults(D=1.0, dt0=1e-3, density=2.65, max_steps=1)
    with pytest.raises(_lib.RockfitError, match="RF_ERR_NODEVICE"):
        fire.relax(w["xyz"], w["r"], c)

"""rockfit -- MI355X (gfx950) bonded-particle DEM step for rock calibration.

Python side of the C-ABI in include/rockfit_dem.h (librockfit_dem.so):
sample generation and I/O, workload configuration, and a Context that owns
one sample's device state.
"""
from . import abi, config, sample  # noqa: F401
from .core import Context, alloc_state, box_reference, derive, state_view  # noqa: F401

__version__ = "0.1.0"

"""FIRE relaxation of a packing on the GPU (rf_fire_relax).

Replaces the relaxation loop of h_CreateCylinderSample
(md_createcylindersample.cu:353-420): 1/rho repulsion of overlapping pairs
(md_calculateforces.cu:20-67), FIRE time-step control (SetFIREData,
md_firerelaxation.cu:72-86), optional cylinder container
(md_cylinderrestriction.cu:69-113) and removal of the worst-overlapping
particle (md_reduceparticlenumber.cu:18-39).  Runs on the GPU only.
"""
import ctypes as C

import numpy as np

from . import abi
from ._lib import check, lib


def defaults(**kw):
    """rf_fire_cfg with the reference's constants, then the given fields."""
    c = abi.FireCfg()
    check(lib().rf_fire_defaults(C.byref(c)), "rf_fire_defaults")
    for k, v in kw.items():
        if k in ("center", "hidden"):
            setattr(c, k, (C.c_double * 3)(*v))
        else:
            setattr(c, k, v)
    return c


def relax(xyz, r, cfg, vel=None, noalloc=False):
    """Relax positions xyz (n, 3) mm with radii r; returns (xyz, vel, result).
    noalloc: use rf_fire_relax_noalloc (working memory in the library's
    static device array, no allocation)."""
    xyz = np.asarray(xyz, dtype=np.float64)
    n = len(xyz)
    x = np.ascontiguousarray(xyz[:, 0]).copy()
    y = np.ascontiguousarray(xyz[:, 1]).copy()
    z = np.ascontiguousarray(xyz[:, 2]).copy()
    rr = np.ascontiguousarray(r, dtype=np.float64)
    v = np.zeros(3 * n) if vel is None else np.ascontiguousarray(np.asarray(vel).T.reshape(-1))
    v = v.astype(np.float64).copy()
    out = abi.FireResult()
    D = C.POINTER(C.c_double)
    fn = lib().rf_fire_relax_noalloc if noalloc else lib().rf_fire_relax
    check(fn(n, x.ctypes.data_as(D), y.ctypes.data_as(D), z.ctypes.data_as(D),
             rr.ctypes.data_as(D), C.byref(cfg), v.ctypes.data_as(D), C.byref(out)),
          "rf_fire_relax_noalloc" if noalloc else "rf_fire_relax")
    return np.stack([x, y, z], axis=1), v.reshape(3, n).T.copy(), out

"""ctypes mirrors of the plain-data structs in include/rockfit_dem.h.

Field order and types must match the header exactly; tests check the sizes
against the C side (``rf_sizeof_*`` are not needed: ctypes computes the same
natural-alignment layout as the C compiler for these all-8-byte/4-byte
members).
"""
import ctypes as C

RF_OK = 0
RF_ERR_ARG = 1
RF_ERR_HIP = 2
RF_ERR_OVERFLOW = 3
RF_ERR_STATE = 4
RF_ERR_GRID = 5
RF_ERR_ALLOC = 6
RF_ERR_IO = 7
RF_ERR_NODEVICE = 8

STATUS_NAMES = {
    0: "RF_OK", 1: "RF_ERR_ARG", 2: "RF_ERR_HIP", 3: "RF_ERR_OVERFLOW", 4: "RF_ERR_STATE",
    5: "RF_ERR_GRID", 6: "RF_ERR_ALLOC", 7: "RF_ERR_IO", 8: "RF_ERR_NODEVICE",
}

RF_QUIRK_BEND_DOUBLE_SCALE = 0x1
RF_QUIRK_ROLL_RESET = 0x2
RF_QUIRK_SLOT0_SENTINEL = 0x4
RF_ORDER_SPATIAL = 0
RF_BRANCH_DECIDE = 0
RF_BRANCH_RECORD = 1
RF_BRANCH_REPLAY = 2
RF_ORDER_GIVEN = 1
RF_QUIRKS_REFERENCE = 0x7

RF_TEST_NONE = 0
RF_TEST_UNIAXIAL = 1
RF_TEST_BRAZILIAN = 2

D3 = C.c_double * 3
I3 = C.c_int32 * 3


class Material(C.Structure):
    _fields_ = [
        ("E_bond", C.c_double), ("nu", C.c_double), ("E_contact", C.c_double),
        ("density", C.c_double), ("restitution", C.c_double), ("friction", C.c_double),
        ("rolling_friction", C.c_double), ("bond_radius", C.c_double),
        ("bond_radius_ratio", C.c_double), ("bond_gap", C.c_double),
        ("sigma_c", C.c_double), ("tau_c", C.c_double),
        ("sigma_c_cv", C.c_double), ("tau_c_cv", C.c_double),
        ("acut", C.c_double), ("dt", C.c_double),
        ("strength_seed", C.c_uint64), ("quirks", C.c_uint32), ("pad_", C.c_uint32),
    ]


class Box(C.Structure):
    _fields_ = [("center", D3), ("half", D3), ("hidden", D3), ("cell", C.c_double),
                ("dims", I3), ("pad_", C.c_int32)]


class Sample(C.Structure):
    _fields_ = [("n", C.c_int64), ("x", C.POINTER(C.c_double)), ("y", C.POINTER(C.c_double)),
                ("z", C.POINTER(C.c_double)), ("r", C.POINTER(C.c_double))]


class Options(C.Structure):
    _fields_ = [("slots", C.c_int32), ("cand_cap", C.c_int32), ("skin", C.c_double),
                ("device", C.c_int32), ("block", C.c_int32), ("stream", C.c_void_p),
                ("order", C.c_int32), ("pad_", C.c_int32)]


class Derived(C.Structure):
    _fields_ = [("G_bond", C.c_double), ("G_contact", C.c_double), ("hm_E", C.c_double),
                ("hm_G", C.c_double), ("p_A", C.c_double), ("T_rayleigh", C.c_double),
                ("T_bond", C.c_double), ("dt", C.c_double), ("r_min", C.c_double),
                ("r_max", C.c_double)]


class PlatenCfg(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("dnsave", C.c_int32), ("v_load", C.c_double),
        ("min_stress", C.c_double), ("max_strain", C.c_double), ("cond_fracture", C.c_double),
        ("size", D3), ("center", D3), ("r_ref", C.c_double), ("wall_modulus_factor", C.c_double),
        ("max_steps", C.c_int64), ("curve_cap", C.c_int64),
    ]


class TestState(C.Structure):
    _fields_ = [
        ("top", C.c_double), ("bottom", C.c_double), ("v", C.c_double),
        ("size_axis", C.c_double), ("size0", C.c_double), ("strain", C.c_double),
        ("stress", C.c_double), ("peak_stress", C.c_double), ("peak_strain", C.c_double),
        ("sfl0", C.c_double), ("sfl1", C.c_double), ("linearity_deviation", C.c_double),
        ("destruction", C.c_double), ("step", C.c_int64), ("steps_stop", C.c_int64),
        ("n_curve", C.c_int64), ("loading", C.c_int32), ("fracture", C.c_int32),
        ("done", C.c_int32), ("kind", C.c_int32),
    ]


class Counters(C.Structure):
    _fields_ = [("n", C.c_int64), ("occupied", C.c_int64), ("bonds", C.c_int64),
                ("contacts", C.c_int64), ("steps", C.c_int64), ("rebuilds", C.c_int64),
                ("candidates", C.c_int64)]


_DP = C.POINTER(C.c_double)


class StateView(C.Structure):
    _fields_ = [
        ("R", _DP), ("V", _DP), ("W", _DP), ("F", _DP), ("M", _DP),
        ("nbr", C.POINTER(C.c_int32)), ("type", C.POINTER(C.c_uint8)),
        ("inv_L0", _DP), ("prev_n", _DP), ("o_t", _DP), ("m_n", _DP), ("m_t", _DP),
        ("sigma_c", _DP), ("tau_c", _DP), ("bidx", C.POINTER(C.c_int8)),
    ]


RF_QUIRK_FIRE_NO_INERTIA = 0x8

# rf_param
RF_PARAM_E_BOND = 0
RF_PARAM_NU = 1
RF_PARAM_E_CONTACT = 2
RF_PARAM_FRICTION = 3
RF_PARAM_ROLLING_FRICTION = 4
RF_PARAM_BOND_RADIUS_RATIO = 5
RF_SENS_MAX = 4
RF_CONTAINER_NONE = 0
RF_CONTAINER_CYLINDER_Z = 1
RF_CONTAINER_CYLINDER_Z_ABSORB = 2
RF_FIRE_LAW_REFERENCE = 0
RF_FIRE_LAW_HARMONIC = 1


class FireCfg(C.Structure):
    _fields_ = [
        ("D", C.c_double), ("dt0", C.c_double), ("density", C.c_double), ("alpha0", C.c_double),
        ("dtgrow", C.c_double), ("dtshrink", C.c_double), ("alpha_shrink", C.c_double),
        ("dtmax_factor", C.c_double), ("dtmin_factor", C.c_double),
        ("ndelay", C.c_int32), ("container", C.c_int32),
        ("np_negative_max", C.c_int64), ("max_steps", C.c_int64), ("hard_cap", C.c_int64),
        ("center", D3), ("radius", C.c_double), ("half_height", C.c_double),
        ("hide_z", C.c_double), ("hidden", D3), ("gap_tol", C.c_double),
        ("check_every", C.c_int32), ("delete_after", C.c_int32), ("skin", C.c_double),
        ("quirks", C.c_uint32), ("device", C.c_int32), ("law", C.c_int32), ("pad_", C.c_int32),
    ]


class Moduli(C.Structure):
    _fields_ = [("secant", C.c_double), ("lsq", C.c_double), ("lsq_intercept", C.c_double),
                ("r2", C.c_double), ("points", C.c_int64)]


class Result(C.Structure):
    _fields_ = [("peak_stress", C.c_double), ("peak_strain", C.c_double),
                ("secant_modulus", C.c_double), ("linearity_deviation", C.c_double),
                ("destruction", C.c_double), ("steps", C.c_int64), ("curve_points", C.c_int64),
                ("fracture", C.c_int32), ("kind", C.c_int32)]


class AdjointCfg(C.Structure):
    _fields_ = [("nparam", C.c_int32), ("params", C.c_int32 * 4), ("pad_", C.c_int32),
                ("steps", C.c_int64)]


class FireResult(C.Structure):
    _fields_ = [
        ("steps", C.c_int64), ("np_positive", C.c_int64), ("np_negative", C.c_int64),
        ("removed", C.c_int64), ("dt", C.c_double), ("alpha", C.c_double),
        ("min_gap", C.c_double), ("power", C.c_double), ("stop_reason", C.c_int32),
        ("pad_", C.c_int32),
    ]


def struct_to_dict(s):
    out = {}
    for name, _ in s._fields_:
        v = getattr(s, name)
        if hasattr(v, "__len__") and not isinstance(v, (bytes, str)):
            v = list(v)
        out[name] = v
    return out

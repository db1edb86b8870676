"""Shared helpers for the parity tests: workloads and state comparison."""
import numpy as np

PARTICLE_FIELDS = ("R", "V", "W", "F", "M")
BOND_FIELDS = ("inv_L0", "prev_n", "o_t", "m_n", "m_t", "sigma_c", "tau_c")


def max_rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.size == 0:
        return 0.0
    scale = max(np.max(np.abs(a)), np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


def compare_states(g, o, n, S, float_tol=0.0, what=""):
    """Compare the device state `g` with the oracle state `o`.

    Lists (partners, kinds, bond numbering) must be identical.  Floating
    fields must agree within `float_tol` relative to the field's largest
    magnitude (0.0 = bitwise, with -0.0 == +0.0).  Only live history is
    compared: bond history on bond slots, o_t on contact slots."""
    errs = []
    for f in PARTICLE_FIELDS:
        r = max_rel(g[f], o[f])
        if r > float_tol:
            errs.append(f"{f}: max rel diff {r:.3e}")
    if not np.array_equal(g["nbr"], o["nbr"]):
        bad = np.nonzero(g["nbr"] != o["nbr"])[0]
        rows = sorted({int(q) // S for q in bad})[:4]
        detail = "; ".join(f"particle {p}: gpu {g['nbr'][p * S:(p + 1) * S].tolist()} "
                           f"oracle {o['nbr'][p * S:(p + 1) * S].tolist()}" for p in rows)
        errs.append(f"nbr differs at {len(bad)} slots: {detail}")
    if not np.array_equal(g["type"], o["type"]):
        bad = np.nonzero(g["type"] != o["type"])[0]
        errs.append(f"type differs at {len(bad)} slots")
    bond = (o["type"] == 1) & (o["nbr"] >= 0)
    cont = (o["type"] == 0) & (o["nbr"] >= 0)
    if not np.array_equal(g["bidx"][bond], o["bidx"][bond]):
        errs.append("bond numbering differs")
    for f in BOND_FIELDS:
        comps = 3 if f in ("prev_n", "o_t", "m_n", "m_t") else 1
        gm = g[f].reshape(n * S, comps)[bond]
        om = o[f].reshape(n * S, comps)[bond]
        r = max_rel(gm, om)
        if r > float_tol:
            errs.append(f"bond {f}: max rel diff {r:.3e}")
    gm = g["o_t"].reshape(n * S, 3)[cont]
    om = o["o_t"].reshape(n * S, 3)[cont]
    r = max_rel(gm, om)
    if r > float_tol:
        errs.append(f"contact o_t: max rel diff {r:.3e}")
    if errs:
        raise AssertionError(f"{what} state mismatch:\n  " + "\n  ".join(errs))


TEST_FIELDS_FLOAT = ("top", "bottom", "v", "size_axis", "size0", "strain", "stress",
                     "peak_stress", "peak_strain", "sfl0", "sfl1")
TEST_FIELDS_INT = ("step", "steps_stop", "n_curve", "loading", "fracture", "done")


def compare_test_states(g, o, tol=0.0, what=""):
    errs = []
    for f in TEST_FIELDS_FLOAT:
        a, b = getattr(g, f), getattr(o, f)
        scale = max(abs(a), abs(b), 1e-300)
        if abs(a - b) / scale > tol:
            errs.append(f"{f}: gpu {a!r} oracle {b!r}")
    for f in TEST_FIELDS_INT:
        if getattr(g, f) != getattr(o, f):
            errs.append(f"{f}: gpu {getattr(g, f)} oracle {getattr(o, f)}")
    if errs:
        raise AssertionError(f"{what} controller mismatch:\n  " + "\n  ".join(errs))


def randomize_motion(st, n, seed, vscale, wscale):
    """Give every particle a random velocity and spin (same arrays to both
    sides) so contacts form and break and new rows are appended."""
    rng = np.random.default_rng(seed)
    st["V"][:] = rng.normal(0.0, vscale, 3 * n)
    st["W"][:] = rng.normal(0.0, wscale, 3 * n)
    return st

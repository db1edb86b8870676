"""Mutation check of the history pins (tests/test_oracle_history.py) and of
the lattice macro pin (tests/test_oracle_macro.py).

Each mutant is a copy of oracle/dem_oracle.c with one plausible slip in the
history transport of md_calculateforcesdem3.cu:57-101, 113-118, 138-145,
183-186, 264-279: the sign of one term of one of the nine entries of M
flipped (24 mutants), M transposed, t or the factor 1/2 of phi wrong, the
previous normal not stored, or, for contacts, the projection or the rescale
dropped or the rescale with a wrong index.  Four more carry a slip in the
macro response: the stress area without r_ref (md_data_types.h:212-242), a
platen moving a full instead of a half step (:347-379), a doubled bond
cross-section (md_calculateforcesdem3.cu:20-200), a halved wall modulus
(md_calculationuniaxialcompression.cu:278).  Every mutant is compiled with
gcc and must fail its module; the unmutated source, compiled the same way,
must pass both.  CPU only.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
SRC = os.path.join(ORACLE, "dem_oracle.c")

BOND = "static void bond_row("
CONTACT = "static void contact_row("

# (name, anchor, old, new): replace the first `old` after `anchor`.
_M_TERMS = [
    ("M1x = 1.0 + tz * pz + ty * py", ["M1x = 1.0 - tz * pz + ty * py",
                                      "M1x = 1.0 + tz * pz - ty * py"]),
    ("M1y = pz - tz - ty * px", ["M1y = -pz - tz - ty * px", "M1y = pz + tz - ty * px",
                                 "M1y = pz - tz + ty * px"]),
    ("M1z = -py - tz * px + ty", ["M1z = py - tz * px + ty", "M1z = -py + tz * px + ty",
                                  "M1z = -py - tz * px - ty"]),
    ("M2x = tz - pz - tx * py", ["M2x = -tz - pz - tx * py", "M2x = tz + pz - tx * py",
                                 "M2x = tz - pz + tx * py"]),
    ("M2y = tz * pz + 1.0 + tx * px", ["M2y = -tz * pz + 1.0 + tx * px",
                                       "M2y = tz * pz + 1.0 - tx * px"]),
    ("M2z = -tz * py + px - tx", ["M2z = tz * py + px - tx", "M2z = -tz * py - px - tx",
                                  "M2z = -tz * py + px + tx"]),
    ("M3x = -ty - tx * pz + py", ["M3x = ty - tx * pz + py", "M3x = -ty + tx * pz + py",
                                  "M3x = -ty - tx * pz - py"]),
    ("M3y = -ty * pz + tx - px", ["M3y = ty * pz + tx - px", "M3y = -ty * pz - tx - px",
                                  "M3y = -ty * pz + tx + px"]),
    ("M3z = ty * py + tx * px + 1.0", ["M3z = -ty * py + tx * px + 1.0",
                                       "M3z = ty * py - tx * px + 1.0"]),
]

MUTANTS = []
for _old, _news in _M_TERMS:
    for _q, _new in enumerate(_news):
        MUTANTS.append((f"{_old.split(' =')[0]}_term{_q}", BOND, _old, _new))
MUTANTS += [
    ("M_transposed", BOND, "        /* torsion */",
     "        { double s_; s_ = M1y; M1y = M2x; M2x = s_; s_ = M1z; M1z = M3x; M3x = s_;"
     " s_ = M2z; M2z = M3y; M3y = s_; }\n        /* torsion */"),
    ("t_reversed", BOND, "double tx = pn[1] * nz - pn[2] * ny;",
     "double tx = pn[2] * ny - pn[1] * nz;"),
    ("phi_no_half", BOND, "* nz) * dt * 0.5;", "* nz) * dt;"),
    ("prev_normal_not_stored", BOND, "pn[0] = nx; pn[1] = ny; pn[2] = nz;", ""),
    ("contact_no_projection", CONTACT,
     "double ox = ot[0] - nx * c1, oy = ot[1] - ny * c1, oz = ot[2] - nz * c1;",
     "double ox = ot[0], oy = ot[1], oz = ot[2];"),
    ("contact_no_rescale", CONTACT, "ox *= c1; oy *= c1; oz *= c1;", ""),
    ("contact_rescale_index", CONTACT, "c1 *= ot[0] * ot[0] + ot[1] * ot[1]",
     "c1 *= ot[0] * ot[1] + ot[1] * ot[1]"),
]


# Slips in the macro response, run against tests/test_oracle_macro.py.
MACRO_MUTANTS = [
    ("area_without_r_ref", "int ora_set_test(",
     "c->area = PI * (hx + cfg->r_ref) * (hx + cfg->r_ref);", "c->area = PI * hx * hx;"),
    ("platen_full_step", "static void controller(", "c->top += 0.5 * c->vw * dt;",
     "c->top += c->vw * dt;"),
    ("bond_section_doubled", BOND, "cefn *= b * b * PI;", "cefn *= 2.0 * b * b * PI;"),
    ("wall_modulus_halved", "double Ew = c->cfg.wall_modulus_factor",
     "c->cfg.wall_modulus_factor * c->der.hm_E", "c->der.hm_E"),
]


def _mutate(src, anchor, old, new):
    a = src.index(anchor)
    q = src.index(old, a)
    return src[:q] + new + src[q + len(old):]


def _build(tmp, name, text):
    path = os.path.join(tmp, f"{name}.c")
    with open(path, "w") as f:
        f.write(text)
    so = os.path.join(tmp, f"lib{name}.so")
    subprocess.run(["gcc", "-O1", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-fopenmp",
                    "-std=c11", "-shared", "-I" + ORACLE, "-o", so, path,
                    os.path.join(ORACLE, "fire_oracle.c"), "-lm"],
                   check=True, capture_output=True)
    return so


def _pins_pass(so, module="test_oracle_history.py"):
    env = dict(os.environ, RF_ORACLE_LIB=so)
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", module)],
                       cwd=ROOT, env=env, capture_output=True, text=True)
    return r.returncode == 0


def test_every_mutant_fails_the_history_pins(tmp_path):
    src = open(SRC).read()
    jobs = [("control", src)]
    for name, anchor, old, new in MUTANTS:
        text = _mutate(src, anchor, old, new)
        assert text != src, name
        jobs.append((name, text))
    assert len(MUTANTS) == 31

    def run(job):
        name, text = job
        return name, _pins_pass(_build(str(tmp_path), name, text))

    with ThreadPoolExecutor(max_workers=4) as ex:
        results = dict(ex.map(run, jobs))
    assert results.pop("control"), "the unmutated oracle must pass the history pins"
    survivors = [name for name, passed in results.items() if passed]
    assert not survivors, f"mutants not caught by the history pins: {survivors}"


def test_every_macro_mutant_fails_the_lattice_pin(tmp_path):
    src = open(SRC).read()
    jobs = [("control", src)]
    for name, anchor, old, new in MACRO_MUTANTS:
        text = _mutate(src, anchor, old, new)
        assert text != src, name
        jobs.append((name, text))

    def run(job):
        name, text = job
        return name, _pins_pass(_build(str(tmp_path), name, text), "test_oracle_macro.py")

    with ThreadPoolExecutor(max_workers=4) as ex:
        results = dict(ex.map(run, jobs))
    assert results.pop("control"), "the unmutated oracle must pass the lattice pin"
    survivors = [name for name, passed in results.items() if passed]
    assert not survivors, f"mutants not caught by the lattice pin: {survivors}"


@pytest.mark.parametrize("name,anchor,old,new", MUTANTS[:1] + MACRO_MUTANTS)
def test_mutation_applies(name, anchor, old, new):
    src = open(SRC).read()
    assert old in src[src.index(anchor):]

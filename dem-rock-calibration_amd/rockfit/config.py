"""Workload builders: material, space box and platen configuration.

Units are the reference's reduced units (md_phys_constants.h:5-14): mm, mg,
us; stress 1e9 Pa, velocity 1e3 m/s, density 1e3 kg/m^3.

`reference_material()` restates SetInitialParameters
(mp_determineconjgradient.cu:15-47): the reference's own monodisperse
granite-like parameters.  `material()` builds the polydisperse workloads of
the uniaxial/Brazilian/batch/adjoint configurations (bond E, contact E,
lognormal strengths, bond radius ratio).
"""
import math
import os

import numpy as np

from . import abi, sample

STRESS = 1e9     # Pa per reduced stress unit
LENGTH = 1e-3    # m per mm
VELOCITY = 1e3   # m/s per reduced velocity unit
DENSITY = 1e3    # kg/m^3 per reduced density unit


def box_reference(space, cell):
    """sample_data / cell_data set-up (md_data_types.h:42-55, 87-124) in FP64;
    identical arithmetic to rf_box_reference."""
    b = abi.Box()
    A = 1.01 * cell
    B = [A + s for s in space]
    for k in range(3):
        b.center[k] = 0.5 * (A + B[k])
        b.half[k] = 0.5 * space[k]
    b.hidden[0] = A + 0.5 * space[0]
    b.hidden[1] = A + 0.5 * space[1]
    b.hidden[2] = (math.floor(B[2] / cell) + 3 - 0.5) * cell
    b.cell = cell
    inv = 1.0 / cell
    b.dims[0] = int(math.floor(B[0] * inv)) + 1
    b.dims[1] = int(math.floor(B[1] * inv)) + 1
    b.dims[2] = int(math.floor(B[2] * inv)) + 4
    return b


def material(r_min, r_mean, E_bond_gpa=20.0, E_contact_gpa=50.0, nu=0.25, density=2650.0,
             restitution=0.6, friction=0.5, rolling_friction=0.05, bond_radius_ratio=1.0,
             gap_factor=0.05, sigma_mpa=40.0, tau_mpa=60.0, cv=0.2, seed=20261203,
             acut_factor=3.2, quirks=abi.RF_QUIRKS_REFERENCE, dt=0.0):
    """Polydisperse BPM material of the uniaxial/Brazilian workloads.

    acut = acut_factor * r_mean (the reference's 3.2 p_r with p_r the mean
    radius) and bond gap = gap_factor * r_min."""
    m = abi.Material()
    m.E_bond = E_bond_gpa * 1e9 / STRESS
    m.nu = nu
    m.E_contact = E_contact_gpa * 1e9 / STRESS
    m.density = density / DENSITY
    m.restitution = restitution
    m.friction = friction
    m.rolling_friction = rolling_friction
    m.bond_radius = 0.0
    m.bond_radius_ratio = bond_radius_ratio
    m.bond_gap = gap_factor * r_min
    m.sigma_c = sigma_mpa * 1e6 / STRESS
    m.tau_c = tau_mpa * 1e6 / STRESS
    m.sigma_c_cv = cv
    m.tau_c_cv = cv
    m.acut = acut_factor * r_mean
    m.dt = dt
    m.strength_seed = seed
    m.quirks = quirks
    return m


def reference_material(E=17.1e9, sigma_c=618e6, tau_c=503e6, bond_radius_m=0.4e-3,
                       gap_m=0.05e-3, quirks=abi.RF_QUIRKS_REFERENCE):
    """SetInitialParameters (mp_determineconjgradient.cu:15-47): p_r = 0.5 mm,
    nu = 0.247, rho = 2610, e = 0.6, mu = 0.45, mu_r = 0.05, acut = 3.2 p_r,
    one modulus for bonds and contacts, absolute bond radius, constant
    strengths.  Defaults for the M(...) values are the commented granite set
    of mp_calculateonesample.cu:63-67."""
    p_r = 0.5
    m = abi.Material()
    m.E_bond = E / STRESS
    m.nu = 0.247
    m.E_contact = m.E_bond
    m.density = 2610.0 / DENSITY
    m.restitution = 0.6
    m.friction = 0.45
    m.rolling_friction = 0.05
    m.bond_radius = bond_radius_m / LENGTH
    m.bond_radius_ratio = 0.0
    m.bond_gap = gap_m / LENGTH
    m.sigma_c = sigma_c / STRESS
    m.tau_c = tau_c / STRESS
    m.sigma_c_cv = 0.0
    m.tau_c_cv = 0.0
    m.acut = 3.2 * p_r
    m.dt = 0.0
    m.strength_seed = 0
    m.quirks = quirks
    return m


def place_in_box(xyz, r, mat, space_factor=(1.6, 1.6, 1.2)):
    """Centre the specimen in a reference space box (SpaceSize/LSi ratios of
    mp_calculateonesample.cu:91) with cell edge 2*acut.  Returns (xyz, box)."""
    lo, hi = xyz.min(axis=0), xyz.max(axis=0)
    size = hi - lo + 2.0 * float(r.max())
    cell = 2.0 * mat.acut
    # at least two cells of margin on each side, so small specimens keep their
    # neighbour stencil inside the grid
    space = [float(max(size[k] * space_factor[k], size[k] + 4.0 * cell)) for k in range(3)]
    box = box_reference(space, cell)
    shift = np.array(box.center[:]) - 0.5 * (lo + hi)
    return xyz + shift, box


def platen_cfg(xyz, r, kind, v_load_mps, min_stress_mpa, max_strain, cond_fracture, dnsave,
               wall_modulus_factor, max_steps=0, curve_cap=0):
    """compression_data parameters from the loaded sample: size_real and
    center_real as h_LoadCylinderSample (md_createcylindersample.cu:448-465)
    with r_max as the particle radius."""
    size, center = sample.cylinder_geometry(xyz, r)
    c = abi.PlatenCfg()
    c.kind = kind
    c.dnsave = dnsave
    c.v_load = v_load_mps / VELOCITY
    c.min_stress = min_stress_mpa * 1e6 / STRESS
    c.max_strain = max_strain
    c.cond_fracture = cond_fracture
    for k in range(3):
        c.size[k] = float(size[k])
        c.center[k] = float(center[k])
    c.r_ref = float(r.max())
    c.wall_modulus_factor = wall_modulus_factor
    c.max_steps = max_steps
    c.curve_cap = curve_cap
    return c


def uniaxial_cfg(xyz, r, v_load_mps=-0.1, min_stress_mpa=10.0, max_strain=0.01,
                 cond_fracture=0.1, dnsave=100, max_steps=0, curve_cap=0):
    """Uniaxial driver values (mp_calculateonesample.cu:120-123, dNSaveU = 100);
    wall moduli 2*hm_E, 2*hm_G (md_calculationuniaxialcompression.cu:278)."""
    return platen_cfg(xyz, r, abi.RF_TEST_UNIAXIAL, v_load_mps, min_stress_mpa, max_strain,
                      cond_fracture, dnsave, 2.0, max_steps, curve_cap)


def brazilian_cfg(xyz, r, v_load_mps=-0.1, min_stress_mpa=2.1, max_strain=0.03,
                  cond_fracture=0.1, dnsave=1000, max_steps=0, curve_cap=0):
    """Brazilian driver values (mp_calculateonesample.cu:121-124: BVelocity =
    -5 * 0.02 m/s, dNSaveB = 1000);
    wall moduli hm_E, hm_G (md_calculationbraziltest.cu:198)."""
    return platen_cfg(xyz, r, abi.RF_TEST_BRAZILIAN, v_load_mps, min_stress_mpa, max_strain,
                      cond_fracture, dnsave, 1.0, max_steps, curve_cap)


# Named workloads (sizes as stated; geometry scaled to hold N spheres at the
# generator's packing fraction).
WORKLOADS = {
    "C1": dict(n=5000, r_min=0.20, r_max=0.30, shape="cylinder", aspect=2.0, test="uniaxial"),
    "C2": dict(n=100000, r_min=0.20, r_max=0.35, shape="cylinder", aspect=2.0, test="uniaxial"),
    "C3": dict(n=50000, r_min=0.20, r_max=0.30, shape="disc", aspect=0.5, test="brazilian"),
}


def make_workload(name=None, n=None, r_min=0.2, r_max=0.3, shape="cylinder", aspect=2.0,
                  test="uniaxial", seed=20261203, gen_kwargs=None, mat_kwargs=None,
                  test_kwargs=None, generator="fcc", cache=None):
    """Generate a sample and build everything a test run needs.

    generator "rsi": random sequential insertion then FIRE on the GPU
    (sample.generate_rsi), the loading configurations' generator; "fcc":
    the CPU jittered-lattice generator (sample.generate).  cache: an .npz
    path that holds the generator's output (written on first use, read
    after), so runs of one workload in separate processes skip the
    generation."""
    if name is not None:
        w = dict(WORKLOADS[name])
        n = n if n is not None else w["n"]
        r_min, r_max, shape, aspect, test = (w["r_min"], w["r_max"], w["shape"], w["aspect"],
                                             w["test"])
    if cache is not None and os.path.exists(cache):
        with np.load(cache) as z:
            xyz, r = z["xyz"].copy(), z["r"].copy()
    elif generator == "rsi":
        gk = dict(packing=0.6)
        gk.update(gen_kwargs or {})
        xyz, r, _ = sample.generate_rsi(n, r_min, r_max, shape=shape, aspect=aspect, seed=seed,
                                        **gk)
    else:
        gk = dict(packing=0.6, tol=0.02, jitter=0.05, max_iter=300)
        gk.update(gen_kwargs or {})
        xyz, r = sample.generate(n, r_min, r_max, shape=shape, aspect=aspect, seed=seed, **gk)
    if cache is not None and not os.path.exists(cache):
        np.savez(cache, xyz=xyz, r=r)
    mat = material(r_min, 0.5 * (r_min + r_max), **(mat_kwargs or {}))
    factor = (1.6, 1.6, 1.2) if shape == "cylinder" else (1.3, 1.3, 1.5)
    xyz, box = place_in_box(xyz, r, mat, factor)
    tk = dict(test_kwargs or {})
    if test == "uniaxial":
        cfg = uniaxial_cfg(xyz, r, **tk)
    elif test == "brazilian":
        cfg = brazilian_cfg(xyz, r, **tk)
    else:
        cfg = None
    return dict(xyz=xyz, r=r, mat=mat, box=box, cfg=cfg, n=n, shape=shape)


# ------------------------------------------------- the loading configurations
# C1-C3: one loading test each; C4: a batch of samples; C5: the elastic
# pre-peak window the gradients are taken over.  Every sample comes from the
# random-sequential-insertion + FIRE generator (sample.generate_rsi, GPU).
STRAIN_RATE = 0.1        # 1/s, uniaxial configurations
BRAZILIAN_SPEED = 0.02   # m/s, platens of the Brazilian configuration
SEED = 20261203


def config_workload(name, n=None, seed=None, dnsave=100, cache=None, **test_kwargs):
    """C1 (5k uniaxial), C2 (100k uniaxial) or C3 (50k Brazilian disc) with
    the configuration's loading: platens closing at strain rate 0.1/s times
    the specimen height (uniaxial) or at 0.02 m/s (Brazilian)."""
    idx = {"C1": 1, "C2": 2, "C3": 3}[name]
    w = make_workload(name, n=n, seed=SEED + idx if seed is None else seed, generator="rsi",
                      cache=cache)
    if WORKLOADS[name]["test"] == "uniaxial":
        v = -STRAIN_RATE * w["cfg"].size[2] * 1e-3
        w["cfg"] = uniaxial_cfg(w["xyz"], w["r"], v_load_mps=v, dnsave=dnsave, **test_kwargs)
    else:
        w["cfg"] = brazilian_cfg(w["xyz"], w["r"], v_load_mps=-BRAZILIAN_SPEED, dnsave=dnsave,
                                 **test_kwargs)
    w["name"] = name
    return w


def batch_workloads(count=64, seed=SEED + 4, n_lo=2000, n_hi=40000, cache_dir=None):
    """C4: `count` uniaxial samples, N log-uniform in [n_lo, n_hi], radii
    U[0.20, 0.30] mm, micro-parameters on a Latin hypercube: bond E in
    [5, 50] GPa, normal strength in [10, 100] MPa (shear strength 1.5 times
    it), friction in [0.3, 0.8]; loaded at strain rate 0.1/s.  cache_dir:
    a directory holding each sample's generator output (make_workload's
    cache), so later runs skip the generation."""
    if cache_dir is not None:
        os.makedirs(cache_dir, exist_ok=True)
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(n_lo), np.log(n_hi), count)).astype(int)

    def lhs(lo, hi):
        u = (rng.permutation(count) + rng.uniform(0.0, 1.0, count)) / count
        return lo + (hi - lo) * u

    E, sig, mu = lhs(5.0, 50.0), lhs(10.0, 100.0), lhs(0.3, 0.8)
    out = []
    for k in range(count):
        w = make_workload(n=int(ns[k]), r_min=0.20, r_max=0.30, test="uniaxial",
                          seed=seed + 100 + k, generator="rsi",
                          cache=(None if cache_dir is None else
                                 os.path.join(cache_dir, f"batch_{seed}_{k}_{int(ns[k])}.npz")),
                          mat_kwargs=dict(E_bond_gpa=float(E[k]), sigma_mpa=float(sig[k]),
                                          tau_mpa=float(1.5 * sig[k]), friction=float(mu[k])))
        v = -STRAIN_RATE * w["cfg"].size[2] * 1e-3
        w["cfg"] = uniaxial_cfg(w["xyz"], w["r"], v_load_mps=v, dnsave=100)
        w["lhs"] = dict(E_bond_gpa=float(E[k]), sigma_mpa=float(sig[k]), friction=float(mu[k]))
        out.append(w)
    return out


def gradient_workload(n=20000, steps=2000, seed=SEED + 5, dnsave=50, min_stress_mpa=0.1):
    """C5: a 20k-sphere cylinder (r ~ U[0.20, 0.30] mm) loaded for `steps`
    elastic pre-peak steps to 0.1 % axial strain (closing speed
    0.001 H / (steps dt)), with dt fixed at the value rf_derive gives (the
    parameters the gradients are taken for do not move it)."""
    from . import core

    w = make_workload(n=n, r_min=0.20, r_max=0.30, test="uniaxial", seed=seed, generator="rsi",
                      test_kwargs=dict(dnsave=dnsave, max_steps=100 * steps, max_strain=0.05,
                                       min_stress_mpa=min_stress_mpa))
    m = abi.Material.from_buffer_copy(w["mat"])
    m.dt = core.derive(w["mat"], float(np.min(w["r"])), float(np.max(w["r"]))).dt
    w["mat"] = m
    w["cfg"].v_load = -0.001 * w["cfg"].size[2] / (steps * m.dt)
    w["steps"] = steps
    return w


def workload_hash(w):
    """Hex digest of everything a loading run's result depends on: sphere
    positions and radii, material, box and platen configuration."""
    import hashlib

    h = hashlib.sha256()
    for a in (np.ascontiguousarray(w["xyz"], dtype=np.float64),
              np.ascontiguousarray(w["r"], dtype=np.float64)):
        h.update(a.tobytes())
    for s_ in (w["mat"], w["box"], w["cfg"]):
        if s_ is not None:
            h.update(bytes(s_))
    return h.hexdigest()[:16]

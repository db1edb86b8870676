"""Seeded sample generator and the reference's Sample_*.txt format.

The reference builds its specimens with a time-seeded cuRAND fill followed by
compaction and FIRE (md_createcylindersample.cu:184-439), so its samples
cannot be reproduced.  This module writes its own seeded, polydisperse
specimens instead:

1. radii are drawn uniform in [r_min, r_max];
2. centres start on a jittered face-centred-cubic lattice sized so that the
   solid fraction inside the cylinder (or disc) equals ``packing``;
3. overlaps are relaxed by damped Jacobi pushes (every overlapping pair is
   pushed apart along its normal, the container keeps every sphere inside)
   until the largest overlap is below ``tol`` times r_min.

The result is a dense random packing with many near-touching pairs, which is
what the bond criterion (gap < bond_gap) needs.  It is deterministic for a
given seed on a given machine (numpy PCG64 plus ordered pair sets).

File format (md_createcylindersample.cu:144-182): first line N, then one
``x y z`` line per sphere in metres.  A fourth column, the radius in metres,
is written for polydisperse samples; a three-column file loads with a single
radius.
"""
import numpy as np

LENGTH_CONST = 1e-3  # md_phys_constants.h:5 (mm)


def _fcc_sites(lattice, lo, hi):
    """FCC sites with cube edge `lattice` covering the box [lo, hi]."""
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    n = np.ceil((hi - lo) / lattice).astype(int) + 1
    g = np.stack(np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing="ij"),
                 axis=-1).reshape(-1, 3).astype(np.float64)
    pts = (g[:, None, :] + basis[None, :, :]).reshape(-1, 3) * lattice + lo
    return pts


def _shape_metric(p, shape, R, H):
    """0 at the centre, 1 on the container wall (max-norm of the two axes)."""
    if shape == "cylinder":  # axis z, radius R, height H
        rho = np.hypot(p[:, 0], p[:, 1]) / R
        ax = np.abs(p[:, 2]) / (0.5 * H)
    else:  # disc: axis z (thickness H), radius R in x-y
        rho = np.hypot(p[:, 0], p[:, 1]) / R
        ax = np.abs(p[:, 2]) / (0.5 * H)
    return np.maximum(rho, ax)


def _container_push(p, r, R, H):
    """Move every sphere inside the container |z| <= H/2 - r, rho <= R - r."""
    zmax = 0.5 * H - r
    p[:, 2] = np.clip(p[:, 2], -zmax, zmax)
    rho = np.hypot(p[:, 0], p[:, 1])
    lim = R - r
    over = rho > lim
    if np.any(over):
        s = lim[over] / rho[over]
        p[over, 0] *= s
        p[over, 1] *= s


def _relax_gpu(pts, r, r_min, r_max, R, H, tol, max_iter):
    """Overlap removal by standard FIRE on the GPU: repulsion of overlapping
    pairs, centres held in the cylinder of radius R - r_max and half height
    H/2 - r_max; stops once no pair overlaps by more than tol r_min."""
    from . import abi, fire

    c = fire.defaults(D=1.0, dt0=0.02, density=2.65, quirks=0, max_steps=max_iter,
                      hard_cap=100 * max(max_iter, 1), gap_tol=-0.5 * tol * r_min / r_max,
                      check_every=10, container=abi.RF_CONTAINER_CYLINDER_Z,
                      radius=R - r_max, half_height=0.5 * H - r_max, hide_z=1e30)
    xyz, _, _ = fire.relax(pts, r, c)
    return xyz


def generate(n, r_min, r_max, shape="cylinder", aspect=2.0, packing=0.60, seed=20261203,
             tol=0.02, max_iter=300, jitter=0.05, step=0.5, verbose=False, relax="cpu"):
    """Return (xyz[n,3], r[n]) in mm, centred at the origin.

    shape "cylinder": height = aspect * diameter (axis z).
    shape "disc": thickness = aspect * diameter (axis z), e.g. aspect 0.5.
    relax "cpu": Jacobi overlap relaxation here; "gpu": FIRE relaxation on the
    GPU (rf_fire_relax) inside the cylinder that holds the sphere centres.
    """
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    r = rng.uniform(r_min, r_max, n)
    vs = np.sum(4.0 / 3.0 * np.pi * r ** 3)
    vc = vs / packing
    D = (4.0 * vc / (np.pi * aspect)) ** (1.0 / 3.0)
    R, H = 0.5 * D, aspect * D
    lattice = (4.0 * vc / n) ** (1.0 / 3.0)
    for _ in range(20):
        lo = np.array([-R, -R, -0.5 * H]) - lattice
        hi = -lo
        pts = _fcc_sites(lattice, lo, hi)
        m = _shape_metric(pts, shape, R, H)
        inside = m <= 1.0
        if inside.sum() >= n:
            break
        lattice *= 0.97
    order = np.argsort(m, kind="stable")
    pts = pts[order[:n]].copy()
    pts += rng.uniform(-jitter, jitter, size=pts.shape) * lattice
    perm = rng.permutation(n)
    r = r[perm]
    _container_push(pts, r, R, H)
    if relax == "gpu":
        return _relax_gpu(pts, r, r_min, r_max, R, H, tol, max_iter), r
    skin = 0.3 * r_min
    pairs = None
    ref = None
    rebuilds = 0
    for it in range(max_iter):
        if pairs is None or np.max(np.abs(pts - ref)) > 0.25 * skin:
            tree = cKDTree(pts)
            pairs = tree.query_pairs(2.0 * r_max + skin, output_type="ndarray")
            rebuilds += 1
            ref = pts.copy()
            pi, pj = pairs[:, 0], pairs[:, 1]
            rsum = r[pi] + r[pj]
        d = pts[pj] - pts[pi]
        dist = np.sqrt(np.einsum("ij,ij->i", d, d))
        ov = rsum - dist
        act = ov > 0.0
        worst = ov[act].max() if np.any(act) else 0.0
        if verbose and it % 100 == 0:
            print(f"relax it={it} pairs={len(pairs)} overlapping={act.sum()} worst={worst:.3e}"
                  f" rebuilds={rebuilds}")
        if worst < tol * r_min:
            break
        a_i, a_j = pi[act], pj[act]
        nvec = d[act] / np.maximum(dist[act], 1e-300)[:, None]
        push = (step * 0.5 * ov[act])[:, None] * nvec
        for k in range(3):
            pts[:, k] += (np.bincount(a_j, weights=push[:, k], minlength=n)
                          - np.bincount(a_i, weights=push[:, k], minlength=n))
        _container_push(pts, r, R, H)
    return pts, r


def cylinder_geometry(xyz, r):
    """Specimen size and centre: the box around all sphere surfaces.

    h_LoadCylinderSample (md_createcylindersample.cu:448-465) takes the
    extent of the centres plus 2 p_r; with one radius that is exactly the
    envelope min(x - r) .. max(x + r) used here, which keeps the platens
    touching the outermost spheres for any radius distribution."""
    rr = np.asarray(r, dtype=np.float64)[:, None]
    lo = (xyz - rr).min(axis=0)
    hi = (xyz + rr).max(axis=0)
    size = hi - lo
    center = 0.5 * (hi + lo)
    return size, center


def write_sample(path, xyz_mm, r_mm=None, precision=17):
    """Sample_*.txt: N, then x y z [r] in metres (SaveParticleDATAtxt)."""
    n = len(xyz_mm)
    fmt = "%.{}g".format(precision)
    with open(path, "w") as f:
        f.write(f"{n}\n")
        for i in range(n):
            row = [xyz_mm[i, 0] * LENGTH_CONST, xyz_mm[i, 1] * LENGTH_CONST,
                   xyz_mm[i, 2] * LENGTH_CONST]
            if r_mm is not None:
                row.append(r_mm[i] * LENGTH_CONST)
            f.write(" ".join(fmt % v for v in row) + "\n")


def read_sample(path, default_radius_mm=None):
    """LoadParticleDATAtxt (md_createcylindersample.cu:165-182): x*lc with
    lc = 1/length_const.  Returns (xyz_mm[n,3], r_mm[n] or None)."""
    lc = 1.0 / LENGTH_CONST
    with open(path) as f:
        n = int(f.readline().split()[0])
        rows = [f.readline().split() for _ in range(n)]
    if len(rows) != n or any(len(rw) < 3 for rw in rows):
        raise ValueError(f"{path}: expected {n} rows of x y z [r]")
    arr = np.array([[float(v) for v in rw[:4]] + ([] if len(rw) >= 4 else [np.nan]) for rw in rows])
    xyz = arr[:, :3] * lc
    if np.all(np.isfinite(arr[:, 3])):
        r = arr[:, 3] * lc
    elif default_radius_mm is not None:
        r = np.full(n, float(default_radius_mm))
    else:
        r = None
    return xyz, r

"""Seeded sample generators and the reference's Sample_*.txt format.

The reference builds its specimens with a time-seeded cuRAND fill followed by
compaction and FIRE (md_createcylindersample.cu:184-439), so its samples
cannot be reproduced.  This module writes its own seeded, polydisperse
specimens instead, in one of two ways:

``generate_rsi`` (the loading configurations: random sequential insertion
then FIRE): radii uniform in [r_min, r_max]; spheres inserted one after
another at random centres in the container at 0.6 of their radius, without
overlaps; then grown to full size in stages, each relaxed by FIRE on the GPU
(rf_fire_relax).  Particles come in insertion order, i.e. in no spatial
order.

``generate`` (CPU only, for tests that need no GPU): centres on a jittered
face-centred-cubic lattice sized so that the solid fraction inside the
cylinder (or disc) equals ``packing``, then overlaps relaxed by damped Jacobi
pushes (every overlapping pair pushed apart along its normal, the container
keeping every sphere inside) until the largest overlap is below ``tol``
times r_min.

Either gives a dense random packing with many near-touching pairs, which is
what the bond criterion (gap < bond_gap) needs, deterministic for a given
seed on a given machine.

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


RSI_STAGES = (0.6, 0.75, 0.85, 0.92, 0.97, 1.0)


def _rsi(r, Rc, Hc, rng, batch=4096, tries=16):
    """Random sequential insertion: spheres are inserted in index order, in
    batches, each at uniform random centres in the cylinder of radius Rc and
    half height Hc; a sphere takes the first of its `tries` candidate centres
    that overlaps no sphere already placed, and, inside its batch, no sphere
    of lower index accepted in the same batch.  Spheres left without a place
    try again in the next batch.  No two placed spheres overlap."""
    from scipy.spatial import cKDTree

    n = len(r)
    rmax = float(r.max())
    pos = np.zeros((n, 3))
    done = np.zeros(n, dtype=bool)
    pending = np.arange(n)
    tree, tree_idx = None, None
    rounds = 0
    while pending.size:
        rounds += 1
        if rounds > 100 * (n // batch + 1):
            raise RuntimeError(f"random sequential insertion placed {int(done.sum())} of {n}")
        idx = pending[:batch]
        m = len(idx)
        ang = rng.uniform(0.0, 2.0 * np.pi, (m, tries))
        rad = Rc * np.sqrt(rng.uniform(0.0, 1.0, (m, tries)))
        cand = np.stack([rad * np.cos(ang), rad * np.sin(ang), rng.uniform(-Hc, Hc, (m, tries))],
                        axis=2)
        ok = np.ones((m, tries), dtype=bool)
        if tree is not None:
            dist, nb = tree.query(cand.reshape(-1, 3), k=24, distance_upper_bound=2.0 * rmax)
            have = nb < len(tree_idx)
            rj = np.where(have, r[tree_idx[np.minimum(nb, len(tree_idx) - 1)]], 0.0)
            ri = np.repeat(r[idx], tries)[:, None]
            ok = ~np.any(have & (dist < ri + rj), axis=1).reshape(m, tries)
        has = ok.any(axis=1)
        first = np.argmax(ok, axis=1)
        prop = cand[np.arange(m), first]
        acc = has.copy()
        sel = np.nonzero(has)[0]
        if sel.size > 1:
            pairs = cKDTree(prop[sel]).query_pairs(2.0 * rmax, output_type="ndarray")
            if len(pairs):
                a_, b_ = sel[pairs[:, 0]], sel[pairs[:, 1]]
                lo, hi = np.minimum(a_, b_), np.maximum(a_, b_)
                d = np.linalg.norm(prop[lo] - prop[hi], axis=1)
                clash = d < r[idx[lo]] + r[idx[hi]]
                lo, hi = lo[clash], hi[clash]
                for q in np.argsort(hi, kind="stable"):  # later sphere of each pair, in order
                    if acc[lo[q]] and acc[hi[q]]:
                        acc[hi[q]] = False
        put = idx[acc]
        pos[put] = prop[acc]
        done[put] = True
        pending = np.concatenate([idx[~acc], pending[batch:]])
        tree_idx = np.nonzero(done)[0]
        tree = cKDTree(pos[tree_idx])
    return pos


def generate_rsi(n, r_min, r_max, shape="cylinder", aspect=2.0, packing=0.60, seed=20261203,
                 stages=RSI_STAGES, gap_tol=-0.005, device=0, verbose=False):
    """Seeded random sequential insertion then FIRE relaxation, on the GPU.

    Radii are uniform in [r_min, r_max].  The container (axis z) holds the
    sphere centres in a cylinder of radius D/2 - r_mean and half height
    H/2 - r_mean, with D and H (= aspect D; for a disc, aspect is the
    thickness over the diameter) sized so that the spheres fill `packing` of
    the D x H cylinder.  Spheres are inserted one after another at
    stages[0] times their radius (no overlaps), then grown through the
    stages; each stage is relaxed by FIRE (rf_fire_relax: harmonic overlap
    force, absorbing cylinder walls) until no pair overlaps by more than
    -gap_tol of its radius sum (last stage) or 1e-3 (earlier stages).  The
    reference fills a box at random with a time-seeded generator and relaxes
    it the same way (md_createcylindersample.cu:184-439).  Returns
    (xyz[n, 3], r[n]) in mm in insertion order, centred at the origin, and a
    dict of the stages' FIRE results."""
    from . import abi, fire

    rng = np.random.default_rng(seed)
    r = rng.uniform(r_min, r_max, n)
    rm = float(r.mean())
    vc = float(np.sum(4.0 / 3.0 * np.pi * r ** 3)) / packing
    D = (4.0 * vc / (np.pi * aspect)) ** (1.0 / 3.0)
    H = aspect * D
    Rc, Hc = 0.5 * D - rm, 0.5 * H - rm
    s0 = stages[0]
    pos = _rsi(r * s0, Rc, Hc, rng)
    info = []
    for q, sc in enumerate(stages):
        last = q == len(stages) - 1
        c = fire.defaults(D=1.0, dt0=0.005, density=2.65, quirks=0, max_steps=50,
                          hard_cap=60000 if last else 20000,
                          gap_tol=gap_tol if last else min(gap_tol, -1e-3),
                          check_every=50 if last else 20, delete_after=-1,
                          container=abi.RF_CONTAINER_CYLINDER_Z_ABSORB, center=(0.0, 0.0, 0.0),
                          radius=Rc, half_height=Hc, hide_z=1e30, hidden=(0.0, 0.0, 1e6),
                          law=abi.RF_FIRE_LAW_HARMONIC, device=device)
        pos, _, res = fire.relax(pos, r * sc, c)
        info.append(dict(scale=sc, steps=res.steps, min_gap=res.min_gap, stop=res.stop_reason))
        if verbose:
            print(f"rsi stage {sc}: {res.steps} FIRE steps, min gap {res.min_gap:.3e}, "
                  f"stop {res.stop_reason}", flush=True)
    return pos, r, info


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

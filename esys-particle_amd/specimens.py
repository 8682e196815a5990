"""Seeded synthetic specimens for tests and benchmarks.

Each builder returns a plain dict of numpy arrays and parameters; ``apply(sim, spec)``
feeds it to an ``LsmMpi``-compatible object through the reference's setup calls
(initNeighbourSearch, setSpatialDomain, createParticles, createConnections,
createInteractionGroup, createExclusion, createWall ...), so the same specimen
drives the HIP engine and the CPU oracle.

Packing: random sequential insertion (small specimens) or the grow-and-relax
packer of ``geometry.py`` (large ones); bonds join pairs with
|x_i - x_j| <= 1.01 (r_i + r_j).
"""
import math

import numpy as np

from . import lsm


def _cell_neighbours(pos, cutoff):
    """All pairs (i<j) with |x_i - x_j| <= cutoff (cell-list brute force)."""
    lo = pos.min(axis=0)
    key = np.floor((pos - lo) / cutoff).astype(np.int64)
    dims = key.max(axis=0) + 3
    lin = (key[:, 0] + 1) * dims[1] * dims[2] + (key[:, 1] + 1) * dims[2] + (key[:, 2] + 1)
    order = np.argsort(lin, kind="stable")
    ls = lin[order]
    pairs = []
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                off = dx * dims[1] * dims[2] + dy * dims[2] + dz
                tgt = lin + off
                a = np.searchsorted(ls, tgt, "left")
                b = np.searchsorted(ls, tgt, "right")
                cnt = b - a
                if cnt.sum() == 0:
                    continue
                i = np.repeat(np.arange(len(pos)), cnt)
                first = np.repeat(a, cnt)
                within = np.arange(int(cnt.sum())) - np.repeat(np.cumsum(cnt) - cnt, cnt)
                j = order[first + within]
                m = i < j
                pairs.append(np.stack([i[m], j[m]], 1))
    if not pairs:
        return np.zeros((0, 2), np.int64)
    pr = np.concatenate(pairs)
    d = np.linalg.norm(pos[pr[:, 0]] - pos[pr[:, 1]], axis=1)
    return pr[d <= cutoff]


def rsa_packing(n, radii, box_lo, box_hi, seed, max_tries=200000, gap=0.0):
    """Random sequential non-overlapping insertion inside [box_lo + r, box_hi - r]."""
    rng = np.random.default_rng(seed)
    box_lo, box_hi = np.asarray(box_lo, float), np.asarray(box_hi, float)
    rmax = float(np.max(radii))
    cs = 2 * rmax + gap
    dims = np.maximum(((box_hi - box_lo) / cs).astype(int) + 1, 1)
    grid = {}
    pos = np.zeros((n, 3))
    placed = 0
    tries = 0
    while placed < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("random packing too dense: placed %d of %d" % (placed, n))
        r = radii[placed]
        p = box_lo + r + rng.random(3) * (box_hi - box_lo - 2 * r)
        c = tuple(((p - box_lo) / cs).astype(int))
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for j in grid.get((c[0] + dx, c[1] + dy, c[2] + dz), ()):
                        if np.sum((pos[j] - p) ** 2) < (r + radii[j] + gap) ** 2:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if not ok:
            continue
        pos[placed] = p
        grid.setdefault(c, []).append(placed)
        placed += 1
    return pos


def bond_pairs(pos, rad, factor=1.01):
    """Pairs with |x_i - x_j| <= factor (r_i + r_j), i < j."""
    cand = _cell_neighbours(pos, 2 * factor * float(rad.max()))
    d = np.linalg.norm(pos[cand[:, 0]] - pos[cand[:, 1]], axis=1)
    keep = d <= factor * (rad[cand[:, 0]] + rad[cand[:, 1]])
    return cand[keep]


def gravity_settling(n=10000, seed=1, box=(20.0, 40.0, 20.0), rmin=0.2, rmax=0.5, dt=1e-4,
                     k=1.0, wall_k=1.0, alpha=0.2):
    """Non-rotational gravity settling: NRotSphere, radii U[rmin, rmax], random
    sequential insertion, NRotElastic contacts, elastic bottom wall, gravity."""
    rng = np.random.default_rng(seed)
    rad = rng.uniform(rmin, rmax, n)
    pos = rsa_packing(n, rad, (0, 0, 0), box, seed + 1)
    return {
        "ptype": "NRotSphere", "grid": 2 * rmax + alpha, "alpha": alpha, "dt": dt,
        "domain": ((-1.0, -1.0, -1.0), (box[0] + 1.0, box[1] + 1.0, box[2] + 1.0)),
        "ids": np.arange(n, dtype=np.int32), "pos": pos, "rad": rad,
        "density": 1.0,
        "elastic": [("elastic", k)],
        "gravity": (0.0, -9.81, 0.0),
        "walls": [("bottom", (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), wall_k, (0.0, 0.0, 0.0))],
    }


def bonded_compression(n=1000, seed=3, rmin=0.2, rmax=1.0, box=None, phi=0.3, dt=1e-5,
                       kr=1e5, ks=3e4, kt=3e4, kb=3e4, coh_mean=50.0, coh_sd=5.0, mu=0.6,
                       wall_speed=0.05, alpha=0.2, bond_factor=1.01, packing=None):
    """Bonded rotational uniaxial compression: RotSphere, radii log-uniform in
    [rmin, rmax], bonds for |x_i - x_j| <= 1.01 (r_i + r_j) with per-bond
    tensile/shear cohesion ~ N(coh_mean, coh_sd) (max_n = max_b = pi T,
    max_s = max_t = pi C), RotFriction mu on unbonded contacts, top and bottom
    walls approaching at wall_speed.

    ``box`` defaults to a 1:2:1 prism sized for solid fraction ``phi``;
    ``packing`` = (pos, rad) overrides the packing."""
    rng = np.random.default_rng(seed)
    if packing is None:
        rad = np.exp(rng.uniform(math.log(rmin), math.log(rmax), n))
        rad = np.sort(rad)[::-1].copy()  # largest first: insertion stays feasible
        if box is None:
            vs = float(np.sum(4.0 / 3.0 * math.pi * rad ** 3))
            s = (vs / phi / 2.0) ** (1.0 / 3.0)
            box = (s, 2 * s, s)
        pos = rsa_packing(n, rad, (0, 0, 0), box, seed + 1)
    else:
        pos, rad = packing
        n = len(rad)
        if box is None:
            box = tuple(np.ceil(pos.max(axis=0) + rad.max()))
    pr = bond_pairs(pos, rad, bond_factor)
    brng = np.random.default_rng(seed + 2)
    T = brng.normal(coh_mean, coh_sd, len(pr))
    C = brng.normal(coh_mean, coh_sd, len(pr))
    strengths = np.stack([math.pi * T, math.pi * C, math.pi * C, math.pi * T], 1)
    ly = box[1]
    v_step = wall_speed * dt
    return {
        "ptype": "RotSphere", "grid": 2 * rmax + alpha, "alpha": alpha, "dt": dt,
        "domain": ((-1.0, -1.0, -1.0), (box[0] + 1.0, ly + 1.0, box[2] + 1.0)),
        "ids": np.arange(n, dtype=np.int32), "pos": pos, "rad": rad, "density": 1.0,
        "box": tuple(box), "cohesion": coh_mean,
        "bonds": pr.astype(np.int32),
        "bond_prms": dict(name="bonded", normalK=kr, shearK=ks, torsionK=kt, bendingK=kb,
                          tag=0, strengths=strengths),
        "friction": ("friction", kr, mu, mu, ks),
        "exclusion": ("bonded", "friction"),
        "walls": [("bottom", (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), kr, (0.0, v_step, 0.0)),
                  ("top", (0.0, ly, 0.0), (0.0, -1.0, 0.0), kr, (0.0, -v_step, 0.0))],
    }


def dense_bonded_compression(n=100000, seed=3, rmin=0.2, rmax=1.0, phi=0.58, max_sweeps=2000,
                             snap_sweeps=1500, height=None, **kw):
    """bonded_compression on a dense grow-and-relax packing (geometry.py): radii
    log-uniform in [rmin, rmax], a 1:2:1 prism sized for solid fraction phi (or
    a square prism of the given height, so that specimens of a batch share the
    walls)."""
    from . import geometry
    rng = np.random.default_rng(seed)
    rad = np.exp(rng.uniform(math.log(rmin), math.log(rmax), n))
    vs = float(np.sum(4.0 / 3.0 * math.pi * rad ** 3))
    if height is None:
        s = (vs / phi / 2.0) ** (1.0 / 3.0)
        box = (s, 2 * s, s)
    else:
        w = math.sqrt(vs / phi / height)
        box = (w, float(height), w)
    pos, _ = geometry.dense_packing(rad, box, seed + 1, max_sweeps)
    # relax the remaining overlaps so that contacts sit at (near) exact touching
    pos, _ = geometry.snap_bonds(pos, rad, np.zeros((0, 2), np.int32), box, snap_sweeps, 1e-9)
    spec = bonded_compression(seed=seed, rmin=rmin, rmax=rmax, box=box, packing=(pos, rad), **kw)
    spec["box"] = box
    return spec


def lattice_block(nx=8, ny=12, nz=8, r=0.5, jitter=1e-3, rad_spread=0.0, seed=7, dt=1e-4,
                  kr=1e5, ks=3e4, kt=3e4, kb=3e4, coh_mean=50.0, coh_sd=5.0, mu=0.6,
                  wall_speed=0.05, vel_sd=0.0, alpha=0.2, bond_factor=1.01):
    """Bonded rotational block on a jittered simple-cubic lattice (spacing 2r):
    every lattice neighbour pair is bonded, so bonds start near equilibrium.
    Same interaction set-up as bonded_compression (walls, friction, exclusion)."""
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1)
    pos = (g.reshape(-1, 3) + 0.5) * (2 * r) + 1.0
    n = len(pos)
    pos = pos + rng.normal(0, jitter, pos.shape)
    rad = r * (1 + rng.uniform(-rad_spread, rad_spread, n))
    box = (nx * 2 * r + 2.0, ny * 2 * r + 2.0, nz * 2 * r + 2.0)
    spec = bonded_compression(seed=seed, rmax=float(rad.max()), box=box, dt=dt, kr=kr, ks=ks,
                              kt=kt, kb=kb, coh_mean=coh_mean, coh_sd=coh_sd, mu=mu,
                              wall_speed=wall_speed, alpha=alpha, bond_factor=bond_factor,
                              packing=(pos, rad))
    if vel_sd > 0:
        spec["vel"] = rng.normal(0, vel_sd, (n, 3))
    return spec


class ServoWallLoader(lsm.Runnable):
    """Pre-time-step runnable holding a wall group's wall at a constant normal
    force: applyForceToWall(group, force) every step (the servo of the
    reference's ServoWallLoaderRunnable in its shear-cell tutorial, without the
    ramp), from time step ``start`` on."""

    def __init__(self, sim, group, force, start=1):
        self.sim, self.group, self.force, self.start = sim, group, lsm.Vec3(*force), start

    def run(self):
        # the wall groups list their candidates at the first step's update
        if self.sim.getTimeStep() >= self.start:
            self.sim.applyForceToWall(self.group, self.force)


def shear_cell(n=2000000, seed=11, rmin=0.25, rmax=0.75, aspect=(200.0, 40.0, 100.0), phi=0.35,
               box=None, dt=5e-5, kn=1e4, ks=1e4, mu=0.5, wall_speed=0.1, stress=1.0,
               wall_k=1e4, alpha=0.2, density=1.0, packing="rsa", sweeps=400, progress=None):
    """Frictional shear cell, periodic in x: RotSphere, radii U[rmin, rmax],
    packed at solid fraction ``phi`` in a box of the given aspect (or ``box``):
    random sequential insertion (``packing="rsa"``, loose) or grow-and-relax
    (``"dense"``, contacts at small overlaps, ``sweeps`` relaxation sweeps);
    RotFriction kn, ks, mu between all particles; the particles within 2 rmax
    of the bottom (tag 2) and top (tag 3) are bonded to rough walls
    (NRotBondedWallPrms) that slide at -/+wall_speed in x; the top wall is
    servo-loaded to the normal stress ``stress`` over its area."""
    from . import geometry
    rng = np.random.default_rng(seed)
    rad = rng.uniform(rmin, rmax, n)
    rad = np.sort(rad)[::-1].copy()  # largest first: insertion stays feasible
    if box is None:
        vs = float(np.sum(4.0 / 3.0 * math.pi * rad ** 3))
        s = (vs / phi / (aspect[0] * aspect[1] * aspect[2])) ** (1.0 / 3.0)
        box = (aspect[0] * s, aspect[1] * s, aspect[2] * s)
    if packing == "dense":
        pos, _ = geometry.dense_packing(rad, box, seed + 1, sweeps, periodic_x=True,
                                        progress=progress)
        # relax the remaining overlaps: contacts at (near) exact touching
        pos, _ = geometry.snap_bonds(pos, rad, np.zeros((0, 2), np.int32), box, sweeps // 2,
                                     1e-9, periodic_x=True, progress=progress)
    else:
        pos = geometry.rsa_packing(rad, box, seed + 1, periodic_x=True)
    # the ids follow a random permutation of the insertion order
    perm = rng.permutation(n)
    pos, rad = pos[perm], rad[perm]
    tags = np.ones(n, np.int32)
    tags[pos[:, 1] < 2 * rmax] = 2
    tags[pos[:, 1] > box[1] - 2 * rmax] = 3
    v_step = wall_speed * dt
    area = box[0] * box[2]
    return {
        "ptype": "RotSphere", "grid": 2 * rmax + alpha, "alpha": alpha, "dt": dt,
        "domain": ((0.0, -1.0, -1.0), (box[0], box[1] + 1.0, box[2] + 1.0)),
        "circ": [True, False, False], "box": box,
        "ids": np.arange(n, dtype=np.int32), "pos": pos, "rad": rad, "tags": tags,
        "density": density,
        "friction": ("friction", kn, mu, mu, ks),
        "walls": [],
        "bonded_walls": [("bottom", (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), wall_k, 2,
                          (-v_step, 0.0, 0.0)),
                         ("top", (0.0, box[1], 0.0), (0.0, -1.0, 0.0), wall_k, 3,
                          (v_step, 0.0, 0.0))],
        "servo": ("top_bonds", (0.0, -stress * area, 0.0)),
    }


def apply(sim, spec, steps_per_batch=None):
    """Set a specimen up on an LsmMpi-compatible simulation object."""
    sim.initNeighbourSearch(spec["ptype"], spec["grid"], spec["alpha"])
    sim.setTimeStepSize(spec["dt"])
    lo, hi = spec["domain"]
    sim.setSpatialDomain(lsm.BoundingBox(lsm.Vec3(*lo), lsm.Vec3(*hi)),
                         circDimList=spec.get("circ"))
    n = len(spec["ids"])
    sim.createParticleArrays(spec["ids"], spec["pos"], spec["rad"],
                             mass=spec["rad"] ** 3, tag=spec.get("tags", np.zeros(n, np.int32)),
                             vel=spec.get("vel"))
    if spec.get("density") is not None:
        sim.setParticleDensity(0, 0, spec["density"])
    if "bonds" in spec:
        b = spec["bonds"]
        ids = spec["ids"]
        sim.createConnectionArrays(ids[b[:, 0]], ids[b[:, 1]], np.zeros(len(b), np.int32))
        bp = spec["bond_prms"]
        if "sc" in bp:  # BrittleBeamSC bonds: E, nu, cohesion, friction angle, tension cutoff
            sim.createInteractionGroup(lsm.BrittleBeamSCPrms(bp["name"], *bp["sc"], bp["tag"]))
        else:
            sim.createInteractionGroup(lsm.RotBondPrms(
                bp["name"], bp["normalK"], bp["shearK"], bp["torsionK"], bp["bendingK"],
                1.0, 1.0, 1.0, 1.0, bp["tag"], perBondStrengths=bp.get("strengths")))
    for name, k in spec.get("elastic", []):
        sim.createInteractionGroup(lsm.NRotElasticPrms(name, k))
    if "friction" in spec:
        name, kn, mud, mus, kss = spec["friction"]
        sim.createInteractionGroup(lsm.RotFrictionPrms(name, kn, mud, mus, kss,
                                                       rigid=spec.get("friction_rigid", False)))
    if "exclusion" in spec:
        sim.createExclusion(*spec["exclusion"])
    if spec.get("gravity") is not None:
        sim.createInteractionGroup(lsm.GravityPrms("gravity", lsm.Vec3(*spec["gravity"])))
    for name, o, nrm, k, step in spec.get("walls", []):
        sim.createWall(name, lsm.Vec3(*o), lsm.Vec3(*nrm))
        sim.createInteractionGroup(lsm.NRotElasticWallPrms(name + "_wall", name, k))
        if any(step):
            sim.setWallStepMotion(name, lsm.Vec3(*step))
    for name, o, nrm, k, tag, step in spec.get("bonded_walls", []):
        sim.createWall(name, lsm.Vec3(*o), lsm.Vec3(*nrm))
        sim.createInteractionGroup(lsm.NRotBondedWallPrms(name + "_bonds", name, k, tag))
        if any(step):
            sim.setWallStepMotion(name, lsm.Vec3(*step))
    if spec.get("servo") is not None:
        sim.addPreTimeStepRunnable(ServoWallLoader(sim, *spec["servo"]))
    return sim


def _calibration_member(args):
    """One specimen of calibration_batch (a worker-process task)."""
    n, seed, height, sweeps, coh = args
    return dense_bonded_compression(n=n, seed=seed, height=height, max_sweeps=sweeps[0],
                                    snap_sweeps=sweeps[1], coh_mean=coh, coh_sd=0.1 * coh)


def calibration_batch(nspec=256, seed=21, nmin=2000, nmax=30000, height=20.0,
                      sweeps=(400, 300), workers=None, progress=None):
    """The batched calibration sweep: nspec bonded compression specimens with
    particle counts log-uniform in [nmin, nmax], each with its own seeded
    packing (dense, square prisms of one height so that the batch shares the
    walls) and its own parameters theta = (k_n in [5e4, 2e5], k_s = 0.3 k_n,
    cohesion mean in [20, 80], mu in [0.3, 0.8]).  Returns (specs, theta
    [nspec, 4]).  Packings are built in worker processes (one OpenMP thread
    each), `workers` of them (default: the CPU share, at most 16); with one
    worker they are built here (OpenMP threads inside each packing) and
    `progress` is called every 16 specimens."""
    import concurrent.futures
    import multiprocessing
    import os
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(math.log(nmin), math.log(nmax), nspec)).astype(int)
    kn = rng.uniform(5e4, 2e5, nspec)
    coh = rng.uniform(20.0, 80.0, nspec)
    mu = rng.uniform(0.3, 0.8, nspec)
    theta = np.stack([kn, 0.3 * kn, coh, mu], 1)
    tasks = [(int(ns[s]), seed * 1000 + s, height, sweeps, float(coh[s])) for s in range(nspec)]
    if workers is None:
        try:
            workers = len(os.sched_getaffinity(0))
        except AttributeError:
            workers = os.cpu_count() or 1
        workers = max(1, min(workers, 16))
    if workers == 1:
        specs = []
        for k, t in enumerate(tasks):
            specs.append(_calibration_member(t))
            if progress and (k + 1) % 16 == 0:
                progress("%d of %d specimens built" % (k + 1, nspec))
        return specs, theta
    old = os.environ.get("OMP_NUM_THREADS")
    os.environ["OMP_NUM_THREADS"] = "1"
    try:
        import runpy
        ctx = multiprocessing.get_context("spawn")
        boot = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_bootstrap.py")
        with concurrent.futures.ProcessPoolExecutor(workers, mp_context=ctx,
                                                    initializer=runpy.run_path,
                                                    initargs=(boot,)) as ex:
            specs = list(ex.map(_calibration_member, tasks, chunksize=1))
    finally:
        if old is None:
            os.environ.pop("OMP_NUM_THREADS", None)
        else:
            os.environ["OMP_NUM_THREADS"] = old
    return specs, theta

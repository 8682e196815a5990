"""Micro-parameter calibration: the differentiable, batched engine
(include/esp_diff.h) behind a torch.autograd.Function.

    spec = specimens.bonded_compression(n=2000, ...)
    model = DiffModel.from_specimens([spec], steps=2000, sample_every=20)
    loss = CalibrationLoss.apply(theta, model, target)   # theta: (nspec, 4)
    loss.sum().backward()                                 # theta.grad

theta columns: (k_n, k_s, cohesion, mu).  k_n is the bonds' and the
contacts' normalK, k_s their shearK, mu the contacts' friction coefficient,
cohesion scales every bond's break values (a specimen's relative break values
are its per-bond strengths divided by its cohesion).  The stress sampled every
`sample_every` steps is the top wall's normal force on the specimen over its
loaded area; loss = mean squared difference to the target curve.

A model holds one or more specimens in one particle set; specimens of a batch
are placed side by side along x with gaps wider than a neighbour-table cell,
and share the walls (one height), so each advances exactly as it would alone.
"""
import ctypes
import os

import numpy as np

from . import _native

HERE = os.path.dirname(os.path.abspath(__file__))

_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int)


class DiffConfig(ctypes.Structure):
    _fields_ = [
        ("grid_spacing", ctypes.c_double),
        ("verlet_dist", ctypes.c_double),
        ("domain_min", ctypes.c_double * 3),
        ("domain_max", ctypes.c_double * 3),
        ("dt", ctypes.c_double),
        ("device", ctypes.c_int),
        ("checkpoint_every", ctypes.c_int),
        ("block", ctypes.c_int),
    ]


_SIGS = {
    "esp_diff_create": (ctypes.c_int, [ctypes.POINTER(DiffConfig), ctypes.POINTER(ctypes.c_void_p)]),
    "esp_diff_destroy": (None, [ctypes.c_void_p]),
    "esp_diff_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "esp_diff_create_error": (ctypes.c_char_p, []),
    "esp_diff_set_particles": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _pi, _pd,
                                              _pd, _pd, _pd]),
    "esp_diff_set_bonds": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _pi, _pi, _pd, _pd, _pd]),
    "esp_diff_add_wall": (ctypes.c_int, [ctypes.c_void_p, _pd, _pd, ctypes.c_double, _pd,
                                         ctypes.c_int, _pd]),
    "esp_diff_set_gravity": (ctypes.c_int, [ctypes.c_void_p, _pd]),
    "esp_diff_set_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "esp_diff_forward": (ctypes.c_int, [ctypes.c_void_p, _pd, _pd, _pd, _pd]),
    "esp_diff_backward": (ctypes.c_int, [ctypes.c_void_p, _pd]),
    "esp_diff_forward_tangent": (ctypes.c_int, [ctypes.c_void_p, _pd, _pd, _pd, _pd, _pd, _pd]),
    "esp_diff_get_particle_data": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _pd]),
    "esp_diff_get_bond_breaks": (ctypes.c_int, [ctypes.c_void_p, _pi, _pi, _pi]),
    "esp_diff_get_rebuild_steps": (ctypes.c_int, [ctypes.c_void_p, _pi]),
    "esp_diff_get_times": (ctypes.c_int, [ctypes.c_void_p, _pd, _pd]),
    "esp_diff_get_branch_digest": (ctypes.c_int, [ctypes.c_void_p,
                                                  ctypes.POINTER(ctypes.c_uint64)]),
    "esp_diff_track_branches": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
}

_LIB = None


def _lib():
    """The HIP engine library with the esp_diff_ entry points (loud if absent)."""
    global _LIB
    if _LIB is None:
        lib = _native.load_hip_library().dll
        for name, (res, args) in _SIGS.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _LIB = lib
    return _LIB


def _dp(a):
    return a.ctypes.data_as(_pd)


def _ip(a):
    return a.ctypes.data_as(_pi)


class DiffModel:
    """One esp_diff handle: specimens, walls, run length and sampling."""

    def __init__(self, grid, alpha, dt, domain, nsteps, sample_every, checkpoint_every=100,
                 device=0, block=256):
        self.lib = _lib()
        cfg = DiffConfig()
        cfg.grid_spacing, cfg.verlet_dist, cfg.dt = float(grid), float(alpha), float(dt)
        for k in range(3):
            cfg.domain_min[k] = float(domain[0][k])
            cfg.domain_max[k] = float(domain[1][k])
        cfg.device, cfg.checkpoint_every, cfg.block = int(device), int(checkpoint_every), int(block)
        h = ctypes.c_void_p()
        rc = self.lib.esp_diff_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise RuntimeError("esp_diff_create failed (%d): %s"
                               % (rc, self.lib.esp_diff_create_error().decode()))
        self.h = h
        self.nsteps, self.every = int(nsteps), int(sample_every)
        self.nsamples = self.nsteps // self.every
        self._check(self.lib.esp_diff_set_run(self.h, self.nsteps, self.every), "set_run")
        self.nspec = 0

    def _check(self, rc, what):
        if rc < 0:
            raise RuntimeError("esp_diff %s failed (%d): %s"
                               % (what, rc, self.lib.esp_diff_last_error(self.h).decode()))
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.lib.esp_diff_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- set-up ----
    def set_particles(self, spec_ids, pos, rad, mass, vel=None):
        spec_ids = np.ascontiguousarray(spec_ids, np.int32)
        pos = np.ascontiguousarray(pos, np.float64)
        rad = np.ascontiguousarray(rad, np.float64)
        mass = np.ascontiguousarray(mass, np.float64)
        v = None if vel is None else np.ascontiguousarray(vel, np.float64)
        self.nspec = int(spec_ids.max()) + 1
        self.n = len(rad)
        self._check(self.lib.esp_diff_set_particles(self.h, len(rad), self.nspec, _ip(spec_ids),
                                                    _dp(pos), _dp(rad), _dp(mass),
                                                    None if v is None else _dp(v)), "set_particles")

    def set_bonds(self, i, j, rel, kt, kb):
        i = np.ascontiguousarray(i, np.int32)
        j = np.ascontiguousarray(j, np.int32)
        rel = np.ascontiguousarray(rel, np.float64).reshape(-1, 4)
        kt = np.ascontiguousarray(kt, np.float64)
        kb = np.ascontiguousarray(kb, np.float64)
        self._check(self.lib.esp_diff_set_bonds(self.h, len(i), _ip(i), _ip(j), _dp(rel), _dp(kt),
                                                _dp(kb)), "set_bonds")

    def add_wall(self, origin, normal, k, step, measure=False, area=None):
        o = np.ascontiguousarray(origin, np.float64)
        nn = np.ascontiguousarray(normal, np.float64)
        st = np.ascontiguousarray(step, np.float64)
        a = None if area is None else np.ascontiguousarray(area, np.float64)
        self._check(self.lib.esp_diff_add_wall(self.h, _dp(o), _dp(nn), float(k), _dp(st),
                                               int(bool(measure)), None if a is None else _dp(a)),
                    "add_wall")

    # ---- runs ----
    def forward(self, theta, target):
        """(stress [nspec, nsamples], loss [nspec]) of one run."""
        th = np.ascontiguousarray(theta, np.float64).reshape(self.nspec, 4)
        tg = np.ascontiguousarray(target, np.float64).reshape(self.nspec, self.nsamples)
        stress = np.zeros((self.nspec, max(self.nsamples, 1)))
        loss = np.zeros(self.nspec)
        self._check(self.lib.esp_diff_forward(self.h, _dp(th), _dp(tg), _dp(stress), _dp(loss)),
                    "forward")
        return stress[:, :self.nsamples], loss

    def forward_tangent(self, theta, dtheta, target):
        """(stress, loss, dloss): a forward run that also carries the state's
        tangent along the parameter direction dtheta [nspec, 4] (forward mode);
        dloss[s] = d loss_s / d eps at theta + eps dtheta."""
        th = np.ascontiguousarray(theta, np.float64).reshape(self.nspec, 4)
        dth = np.ascontiguousarray(dtheta, np.float64).reshape(self.nspec, 4)
        tg = np.ascontiguousarray(target, np.float64).reshape(self.nspec, self.nsamples)
        stress = np.zeros((self.nspec, max(self.nsamples, 1)))
        loss = np.zeros(self.nspec)
        dloss = np.zeros(self.nspec)
        self._check(self.lib.esp_diff_forward_tangent(self.h, _dp(th), _dp(dth), _dp(tg),
                                                      _dp(stress), _dp(loss), _dp(dloss)),
                    "forward_tangent")
        return stress[:, :self.nsamples], loss, dloss

    def backward(self):
        """d loss_s / d theta_s of the last forward run, [nspec, 4]."""
        g = np.zeros((self.nspec, 4))
        self._check(self.lib.esp_diff_backward(self.h, _dp(g)), "backward")
        return g

    def particle_data(self, field):
        f = _native.FIELD_NAMES[field]
        w = 4 if f == _native.FIELD_QUAT else 3
        out = np.zeros((self.n, w))
        self._check(self.lib.esp_diff_get_particle_data(self.h, f, _dp(out)), "get_particle_data")
        return out

    def bond_breaks(self):
        """(step the bond left the group, i, j) rows."""
        m = self._check(self.lib.esp_diff_get_bond_breaks(self.h, None, None, None), "bond_breaks")
        s, a, b = (np.zeros(max(m, 1), np.int32) for _ in range(3))
        self._check(self.lib.esp_diff_get_bond_breaks(self.h, _ip(s), _ip(a), _ip(b)), "bond_breaks")
        return np.stack([s[:m], a[:m], b[:m]], 1)

    def rebuild_steps(self):
        m = self._check(self.lib.esp_diff_get_rebuild_steps(self.h, None), "rebuild_steps")
        s = np.zeros(max(m, 1), np.int32)
        self._check(self.lib.esp_diff_get_rebuild_steps(self.h, _ip(s)), "rebuild_steps")
        return s[:m]

    def track_branches(self, enable=True):
        """Compute the per-step branch digests of later forward runs (off by
        default: it costs forward time)."""
        self._check(self.lib.esp_diff_track_branches(self.h, int(bool(enable))), "track_branches")

    def branch_digest(self):
        """Per-step branch digests of the last forward run (uint64 [nsteps]):
        contact touching/slipping and wall contacts hashed per step."""
        m = self._check(self.lib.esp_diff_get_branch_digest(self.h, None), "branch_digest")
        out = np.zeros(max(m, 1), np.uint64)
        self._check(self.lib.esp_diff_get_branch_digest(
            self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))), "branch_digest")
        return out[:m]

    def times_ms(self):
        f, b = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.esp_diff_get_times(self.h, ctypes.byref(f), ctypes.byref(b)), "times")
        return f.value, b.value

    # ---- specimens ----
    @staticmethod
    def from_specimens(specs, nsteps, sample_every, gap=None, checkpoint_every=100, device=0,
                       block=256, wall_k=None, x_start=0.0, grid=None, alpha=None):
        """A model of bonded compression specimens (specimens.bonded_compression
        dicts: RotSphere, one RotBond group with per-bond strengths, RotFriction,
        top and bottom walls moving at their steps).  Specimen s is shifted by
        x_s along x (the first one's left edge at x_start; model.starts holds
        every specimen's left edge); all must share the walls (same y extent and
        wall motion).  Coordinates round differently at different offsets, so a
        specimen runs bitwise as it would alone only at the same x_s."""
        specs = list(specs)
        s0 = specs[0]
        dt = s0["dt"]
        alpha = s0["alpha"] if alpha is None else float(alpha)
        grid = s0["grid"] if grid is None else float(grid)
        gap = 2.0 * max(grid, s0["grid"]) if gap is None else gap
        pos, rad, mass, sid, bi, bj, rel = [], [], [], [], [], [], []
        kt, kb, area = [], [], []
        x = float(x_start)
        base = 0
        offsets, starts = [], []
        for s, sp in enumerate(specs):
            starts.append(x)
            p = np.array(sp["pos"], np.float64)
            lo = float(p[:, 0].min() - sp["rad"].max())
            shift = x - lo
            offsets.append(shift)
            p = p.copy()
            p[:, 0] += shift
            r = np.asarray(sp["rad"], np.float64)
            # setParticleDensity: m = 4/3 rho pi r^3 with pi = 3.141592654
            # (CParticle::setDensity, Particle.cpp:108-122)
            rho = sp.get("density", 1.0)
            m = (4.0 / 3.0) * rho * 3.141592654 * r * r * r
            pos.append(p)
            rad.append(r)
            mass.append(m)
            sid.append(np.full(len(r), s, np.int32))
            b = np.asarray(sp["bonds"], np.int64)
            coh = sp.get("cohesion", 50.0)
            rel.append(np.asarray(sp["bond_prms"]["strengths"], np.float64) / coh)
            bi.append(b[:, 0] + base)
            bj.append(b[:, 1] + base)
            kt.append(sp["bond_prms"]["torsionK"])
            kb.append(sp["bond_prms"]["bendingK"])
            box = sp["box"] if "box" in sp else sp["domain"][1]
            area.append(float(box[0]) * float(box[2]))
            base += len(r)
            x = float(p[:, 0].max() + r.max()) + gap
        pos, rad, mass, sid = (np.concatenate(a) for a in (pos, rad, mass, sid))
        bi, bj, rel = np.concatenate(bi), np.concatenate(bj), np.concatenate(rel)
        lo = np.minimum(pos.min(0) - 2 * grid, np.array(s0["domain"][0]))
        hi = np.maximum(pos.max(0) + 2 * grid, np.array(s0["domain"][1]))
        model = DiffModel(grid, alpha, dt, (lo, hi), nsteps, sample_every, checkpoint_every,
                          device, block)
        model.set_particles(sid, pos, rad, mass)
        model.set_bonds(bi, bj, rel, kt, kb)
        walls = []
        for name, o, nrm, k, step in s0["walls"]:
            kw = k if wall_k is None else wall_k
            walls.append((name, o, nrm, kw, step))
            model.add_wall(o, nrm, kw, step, measure=(name == "top"),
                           area=np.array(area) if name == "top" else None)
        model.offsets = offsets
        model.starts = starts
        model.specs = specs
        model.host = {"grid": grid, "alpha": alpha, "dt": dt, "domain": (lo, hi), "pos": pos,
                      "rad": rad, "mass": mass, "sid": sid, "bi": bi, "bj": bj, "rel": rel,
                      "kt": np.array(kt, float), "kb": np.array(kb, float), "walls": walls,
                      "area": np.array(area, float)}
        return model


def autotune(specs, theta, trial_steps=60, sample_every=20, blocks=(128, 256, 512),
             cell_factors=(1.0, 1.3), alpha_factors=(0.5, 1.0, 2.0), device=0, log=None):
    """Launch and neighbour-search settings for a batch of specimens, chosen by
    timing short forward runs: the block size of every kernel, the Verlet skin
    alpha (the rebuild interval) and the neighbour-table cell size (a factor on
    2 r_max + alpha).  None of them may change a result: every trial's break
    events and final state must equal the first trial's bit for bit, else the
    setting is rejected.  Coordinate descent: block, then alpha, then cell.
    Returns {"block", "alpha", "grid", "gap" (the specimens' spacing all trials
    used; the model must use it too), "trials": [(block, alpha, grid, ms, ok)]}."""
    specs = list(specs)
    rmax = max(float(np.max(sp["rad"])) for sp in specs)
    a0 = float(specs[0]["alpha"])
    th = np.asarray(theta, np.float64).reshape(len(specs), 4)
    # one placement of the specimens for every trial (coordinates round alike)
    gap = 2.0 * max(cell_factors) * (2.0 * rmax + max(alpha_factors) * a0)
    ref = {}
    trials = []

    def trial(block, alpha, cell):
        grid = cell * (2.0 * rmax + alpha)
        m = DiffModel.from_specimens(specs, trial_steps, sample_every, device=device,
                                     block=block, grid=grid, alpha=alpha, gap=gap)
        tgt = np.zeros((m.nspec, m.nsamples))
        m.forward(th, tgt)  # first run: allocations, code loading
        st, _ = m.forward(th, tgt)
        ms = m.times_ms()[0]
        state = (m.bond_breaks().tobytes(), m.particle_data("pos").tobytes(),
                 m.particle_data("quat").tobytes(), st.tobytes())
        m.close()
        if not ref:
            ref["state"] = state
        ok = state == ref["state"]
        trials.append((block, alpha, grid, ms, ok))
        if log:
            log("autotune: block %d alpha %.3g grid %.4g: %.2f ms%s"
                % (block, alpha, grid, ms, "" if ok else " (results differ: rejected)"))
        return ms if ok else float("inf")

    best = {"block": 256, "alpha": a0, "cell": 1.0}
    cost = trial(best["block"], best["alpha"], best["cell"])
    for key, values in (("block", blocks), ("alpha", [f * a0 for f in alpha_factors]),
                        ("cell", cell_factors)):
        for v in values:
            if v == best[key]:
                continue
            cand = dict(best)
            cand[key] = v
            c = trial(cand["block"], cand["alpha"], cand["cell"])
            if c < cost:
                best, cost = cand, c
    return {"block": best["block"], "alpha": best["alpha"],
            "grid": best["cell"] * (2.0 * rmax + best["alpha"]), "gap": gap, "trial_ms": cost,
            "trials": trials}


def plain_setup(sim, model, theta, spec_index=0):
    """Specimen `spec_index` of `model` on an LsmMpi-compatible engine (the
    plain engine, or the oracle's facade) with parameters theta = (k_n, k_s,
    cohesion, mu): the same particles, masses, bonds with break values
    cohesion x relative values, RotFriction excluded by the bonds, and walls.
    Returns the particle-index offset of the specimen in the model."""
    from . import lsm
    H = model.host
    sel = np.nonzero(H["sid"] == spec_index)[0]
    a, b = int(sel[0]), int(sel[-1]) + 1
    kn, ks, coh, mu = (float(v) for v in theta)
    sim.initNeighbourSearch("RotSphere", H["grid"], H["alpha"])
    sim.setTimeStepSize(H["dt"])
    lo, hi = H["domain"]
    sim.setSpatialDomain(lsm.BoundingBox(lsm.Vec3(*lo), lsm.Vec3(*hi)))
    sim.createParticleArrays(np.arange(b - a), H["pos"][a:b], H["rad"][a:b], mass=H["mass"][a:b])
    m = (H["bi"] >= a) & (H["bi"] < b)
    sim.createConnectionArrays(H["bi"][m] - a, H["bj"][m] - a, np.zeros(int(m.sum()), np.int32))
    sim.createInteractionGroup(lsm.RotBondPrms(
        "bonded", kn, ks, float(H["kt"][spec_index]), float(H["kb"][spec_index]), 1.0, 1.0, 1.0,
        1.0, 0, perBondStrengths=coh * H["rel"][m]))
    sim.createInteractionGroup(lsm.RotFrictionPrms("friction", kn, mu, mu, ks))
    sim.createExclusion("bonded", "friction")
    for name, o, nrm, k, step in H["walls"]:
        sim.createWall(name, lsm.Vec3(*o), lsm.Vec3(*nrm))
        sim.createInteractionGroup(lsm.NRotElasticWallPrms(name + "_wall", name, k))
        if any(step):
            sim.setWallStepMotion(name, lsm.Vec3(*step))
    return a


def plain_stress(sim, model, spec_index=0):
    """The sampled stresses of a plain-engine run of the model's length: the
    top wall's force (the sum of the forces it applies to the particles, as
    CWall accumulates it) along its normal, over the area, at steps every,
    2 every, ..."""
    H = model.host
    nrm = np.array([w for w in H["walls"] if w[0] == "top"][0][2], float)
    hist = sim.engine.wall_force_history("top", model.nsteps)
    idx = np.arange(model.every, model.nsteps + 1, model.every) - 1
    return (hist[idx] @ nrm) / H["area"][spec_index]


def plain_loss(sim_class, model, theta, target, spec_index=0):
    """Loss of specimen `spec_index` run on a fresh plain engine (a sim class:
    LsmMpi, or the oracle's facade class); also its break events."""
    sim = sim_class()
    plain_setup(sim, model, theta, spec_index)
    sim.runTimeSteps(model.nsteps)
    st = plain_stress(sim, model, spec_index)
    tg = np.asarray(target, float).reshape(-1)
    return float(np.mean((st - tg) ** 2)), st, sim.getBondBreaks()


def theta_of(spec):
    """(k_n, k_s, cohesion, mu) of a bonded_compression specimen."""
    bp = spec["bond_prms"]
    return np.array([bp["normalK"], bp["shearK"], spec.get("cohesion", 50.0), spec["friction"][2]])


try:
    import torch

    class CalibrationLoss(torch.autograd.Function):
        """loss[s] of the specimens of `model` for parameters theta [nspec, 4]
        (float64); backward: the hand-written adjoint of the engine."""

        @staticmethod
        def forward(ctx, theta, model, target):
            th = theta.detach().cpu().double().numpy()
            tg = np.asarray(target.detach().cpu().double().numpy() if hasattr(target, "detach")
                            else target, np.float64)
            _, loss = model.forward(th, tg)
            ctx.model = model
            ctx.dev = theta.device
            return torch.as_tensor(loss, dtype=torch.float64, device=theta.device)

        @staticmethod
        def backward(ctx, gout):
            g = ctx.model.backward()
            go = gout.detach().cpu().double().numpy()[:, None]
            return torch.as_tensor(g * go, dtype=torch.float64, device=ctx.dev), None, None
except ImportError:  # torch absent: the numpy interface above still works
    CalibrationLoss = None

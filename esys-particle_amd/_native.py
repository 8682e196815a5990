"""ctypes binding of the engine C-ABI (include/esp.h).

The product path binds ``libesp.so`` (HIP kernels + host orchestration, built in
this directory by ``build.py``) with the ``esp_`` prefix.  The same table also
binds any library exporting the same entry points under another prefix, which is
how the test suite drives the CPU oracle through identical calls.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libesp.so")

ESP_PTYPE_NROT = 0
ESP_PTYPE_ROT = 1

FIELD_POS, FIELD_VEL, FIELD_ANGVEL, FIELD_QUAT = 0, 1, 2, 3
FIELD_FORCE, FIELD_MOMENT, FIELD_RAD, FIELD_MASS, FIELD_OLDPOS = 4, 5, 6, 7, 8
FIELD_WIDTH = {0: 3, 1: 3, 2: 3, 3: 4, 4: 3, 5: 3, 6: 1, 7: 1, 8: 3}
FIELD_NAMES = {"pos": 0, "vel": 1, "angvel": 2, "quat": 3, "force": 4,
               "moment": 5, "rad": 6, "mass": 7, "oldpos": 8}


class EspConfig(ctypes.Structure):
    _fields_ = [
        ("particle_type", ctypes.c_int),
        ("grid_spacing", ctypes.c_double),
        ("verlet_dist", ctypes.c_double),
        ("domain_min", ctypes.c_double * 3),
        ("domain_max", ctypes.c_double * 3),
        ("dt", ctypes.c_double),
        ("do_2d", ctypes.c_int),
        ("device", ctypes.c_int),
        ("steps_per_batch", ctypes.c_int),
        ("circular_x", ctypes.c_int),
    ]


class EspRotBondPrms(ctypes.Structure):
    _fields_ = [
        ("kr", ctypes.c_double), ("ks", ctypes.c_double),
        ("kt", ctypes.c_double), ("kb", ctypes.c_double),
        ("max_n", ctypes.c_double), ("max_s", ctypes.c_double),
        ("max_t", ctypes.c_double), ("max_b", ctypes.c_double),
        ("tag", ctypes.c_int), ("scaling", ctypes.c_int),
        ("meanR_scaling", ctypes.c_int),
        ("truncated", ctypes.c_double),
        ("beta1", ctypes.c_double), ("beta2", ctypes.c_double),
    ]


class EspBrittleBeamSCPrms(ctypes.Structure):
    _fields_ = [
        ("kr", ctypes.c_double), ("ks", ctypes.c_double),
        ("kt", ctypes.c_double), ("kb", ctypes.c_double),
        ("cohesion", ctypes.c_double), ("friction_angle", ctypes.c_double),
        ("tension_cutoff", ctypes.c_double), ("tag", ctypes.c_int),
    ]


class EspRotFrictionPrms(ctypes.Structure):
    _fields_ = [
        ("k_n", ctypes.c_double), ("mu_d", ctypes.c_double),
        ("mu_s", ctypes.c_double), ("k_s", ctypes.c_double),
        ("scaling", ctypes.c_int), ("rigid", ctypes.c_int),
        ("meanR_scaling", ctypes.c_int),
    ]


class EspStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_longlong) for n in (
        "time_step", "num_rebuilds", "num_particles", "num_bonds", "num_dynamic",
        "num_candidates", "num_broken", "num_host_phases")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_s = ctypes.c_char_p
_pi = ctypes.POINTER(ctypes.c_int)
_pd = ctypes.POINTER(ctypes.c_double)

SIGNATURES = {
    "create": (_i, [ctypes.POINTER(EspConfig), ctypes.POINTER(_vp)]),
    "destroy": (None, [_vp]),
    "last_error": (_s, [_vp]),
    "create_error": (_s, []),
    "set_time_step_size": (_i, [_vp, _d]),
    "add_particles": (_i, [_vp, _i, _pi, _pi, _pd, _pd, _pd, _pd]),
    "set_particle_density": (_i, [_vp, _i, _i, _d]),
    "set_particle_non_dynamic": (_i, [_vp, _i]),
    "set_particle_non_rotational": (_i, [_vp, _i]),
    "add_connections": (_i, [_vp, _i, _pi, _pi, _pi]),
    "add_rot_bond_group": (_i, [_vp, _s, ctypes.POINTER(EspRotBondPrms), _pd, _i]),
    "add_brittle_beam_sc_group": (_i, [_vp, _s, ctypes.POINTER(EspBrittleBeamSCPrms)]),
    "add_nrot_elastic_group": (_i, [_vp, _s, _d, _i]),
    "add_rot_elastic_group": (_i, [_vp, _s, _d, _i]),
    "add_rot_friction_group": (_i, [_vp, _s, ctypes.POINTER(EspRotFrictionPrms)]),
    "add_exclusion": (_i, [_vp, _s, _s]),
    "add_wall": (_i, [_vp, _s, _pd, _pd]),
    "add_elastic_wall_group": (_i, [_vp, _s, _s, _d]),
    "add_gravity": (_i, [_vp, _s, _pd]),
    "add_damping_group": (_i, [_vp, _i, _s, ctypes.c_double, _i]),
    "add_bonded_wall_group": (_i, [_vp, _s, _s, ctypes.c_double, _i, _i]),
    "set_interaction_group_tags": (_i, [_vp, _s, _i, _i, _i, _i]),
    "set_wall_group_tags": (_i, [_vp, _s, _i, _i]),
    "set_particle_vel": (_i, [_vp, _i, _pd]),
    "set_particle_ang_vel": (_i, [_vp, _i, _pd]),
    "move_wall_by": (_i, [_vp, _s, _pd]),
    "apply_force_to_wall": (_i, [_vp, _s, _pd]),
    "set_wall_step_motion": (_i, [_vp, _s, _pd]),
    "run_time_steps": (_i, [_vp, _i]),
    "get_num_particles": (_i, [_vp]),
    "get_stats": (_i, [_vp, ctypes.POINTER(EspStats)]),
    "get_time_step": (_i, [_vp, ctypes.POINTER(ctypes.c_longlong)]),
    "get_particle_ids": (_i, [_vp, _pi]),
    "get_particle_data": (_i, [_vp, _i, _pd]),
    "get_wall_position": (_i, [_vp, _s, _pd]),
    "get_wall_force": (_i, [_vp, _s, _pd]),
    "get_wall_force_history": (_i, [_vp, _s, _i, _pd]),
    "get_interaction_count": (_i, [_vp, _s]),
    "get_interactions": (_i, [_vp, _s, _pi, _pi]),
    "get_friction_state": (_i, [_vp, _s, _pd, _pi]),
    "get_neighbour_pairs": (_i, [_vp, _pi, _pi]),
    "get_particle_stress": (_i, [_vp, _pd]),
    "save_checkpoint": (_i, [_vp, _s, _i, _i]),
    "load_checkpoint": (_i, [_vp, _s]),
    "get_bond_breaks": (_i, [_vp, _pi, _pi, _pi]),
    "get_rebuild_steps": (_i, [_vp, _pi]),
    "track_pair_events": (_i, [_vp, _s]),
    "get_pair_events": (_i, [_vp, _s, _pi, _pi, _pi, _pi]),
    "track_bond_criteria": (_i, [_vp]),
    "get_bond_criteria": (_i, [_vp, _s, _pi, _pi, _pd]),
    "set_kernel_timing": (_i, [_vp, _i]),
    "get_kernel_times": (_i, [_vp, _pd, ctypes.POINTER(ctypes.c_longlong)]),
}

KERNEL_CLASSES = ("forces", "gather_integrate")


# entry points of the HIP library only (the oracle has no counterpart)
HIP_ONLY_SIGNATURES = {
    "apply_force_to_wall_noalloc": (_i, [_vp, _s, _pd]),
    "check_state": (_i, [_vp]),
}


class NativeLib:
    """Function table over one shared library exporting the engine C-ABI."""

    def __init__(self, path, prefix, extra=None):
        self.path = path
        self.prefix = prefix
        self.dll = ctypes.CDLL(path)
        table = dict(SIGNATURES)
        if extra:
            table.update(extra)
        for name, (res, args) in table.items():
            fn = getattr(self.dll, prefix + name)
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)


_HIP_LIB = None


def load_hip_library():
    """Load the HIP engine library; raises if it has not been built."""
    global _HIP_LIB
    if _HIP_LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libesp.so is not built (expected %s); run __graft_entry__.build()" % LIB_PATH)
        _HIP_LIB = NativeLib(LIB_PATH, "esp_", HIP_ONLY_SIGNATURES)
    return _HIP_LIB


def _ip(a):
    return a.ctypes.data_as(_pi) if a is not None else None


def _dp(a):
    return a.ctypes.data_as(_pd) if a is not None else None


def _b(s):
    return s.encode() if isinstance(s, str) else s


class Engine:
    """One engine handle (one subdomain) over a NativeLib."""

    def __init__(self, lib, particle_type, grid_spacing, verlet_dist, domain_min,
                 domain_max, dt=0.0, do_2d=False, device=0, steps_per_batch=32,
                 circular_x=False):
        self.lib = lib
        cfg = EspConfig()
        cfg.particle_type = particle_type
        cfg.grid_spacing = float(grid_spacing)
        cfg.verlet_dist = float(verlet_dist)
        for k in range(3):
            cfg.domain_min[k] = float(domain_min[k])
            cfg.domain_max[k] = float(domain_max[k])
        cfg.dt = float(dt)
        cfg.do_2d = int(bool(do_2d))
        cfg.device = int(device)
        cfg.steps_per_batch = int(steps_per_batch)
        cfg.circular_x = int(bool(circular_x))
        h = ctypes.c_void_p()
        rc = lib.create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise RuntimeError("engine create failed (%d): %s" % (rc, lib.create_error().decode()))
        self.h = h
        self.particle_type = particle_type

    def close(self):
        if getattr(self, "h", None):
            self.lib.destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc < 0:
            raise RuntimeError("%s failed (%d): %s" % (what, rc, self.lib.last_error(self.h).decode()))
        return rc

    # ---- setup ----
    def set_time_step_size(self, dt):
        self._check(self.lib.set_time_step_size(self.h, float(dt)), "setTimeStepSize")

    def add_particles(self, ids, pos, rad, mass=None, tags=None, vel=None):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        n = len(ids)
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(n, 3)
        rad = np.ascontiguousarray(rad, dtype=np.float64).reshape(n)
        mass = None if mass is None else np.ascontiguousarray(mass, dtype=np.float64).reshape(n)
        tags = None if tags is None else np.ascontiguousarray(tags, dtype=np.int32).reshape(n)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(n, 3)
        self._check(self.lib.add_particles(self.h, n, _ip(ids), _ip(tags), _dp(pos), _dp(rad),
                                           _dp(mass), _dp(vel)), "createParticles")

    def set_particle_density(self, tag, mask, rho):
        self._check(self.lib.set_particle_density(self.h, int(tag), int(mask), float(rho)),
                    "setParticleDensity")

    def set_particle_non_dynamic(self, tag):
        self._check(self.lib.set_particle_non_dynamic(self.h, int(tag)), "setParticleNonDynamic")

    def set_particle_non_rotational(self, tag):
        self._check(self.lib.set_particle_non_rotational(self.h, int(tag)),
                    "setParticleNonRotational")

    def add_connections(self, id1, id2, tags=None):
        id1 = np.ascontiguousarray(id1, dtype=np.int32)
        id2 = np.ascontiguousarray(id2, dtype=np.int32)
        tags = (np.zeros(len(id1), np.int32) if tags is None
                else np.ascontiguousarray(tags, dtype=np.int32))
        self._check(self.lib.add_connections(self.h, len(id1), _ip(id1), _ip(id2), _ip(tags)),
                    "createConnections")

    def add_rot_bond_group(self, name, prms, per_bond_max=None):
        c = EspRotBondPrms()
        for f, _ in EspRotBondPrms._fields_:
            setattr(c, f, prms[f])
        pb = None
        n = 0
        if per_bond_max is not None:
            pb = np.ascontiguousarray(per_bond_max, dtype=np.float64).reshape(-1, 4)
            n = len(pb)
        self._check(self.lib.add_rot_bond_group(self.h, _b(name), ctypes.byref(c), _dp(pb), n),
                    "createInteractionGroup(RotBondPrms)")

    def add_brittle_beam_sc_group(self, name, prms):
        c = EspBrittleBeamSCPrms()
        for f, _ in EspBrittleBeamSCPrms._fields_:
            setattr(c, f, prms[f])
        self._check(self.lib.add_brittle_beam_sc_group(self.h, _b(name), ctypes.byref(c)),
                    "createInteractionGroup(BrittleBeamSCPrms)")

    def save_checkpoint(self, prefix, num_time_steps=0, precision=12):
        self._check(self.lib.save_checkpoint(self.h, _b(prefix), int(num_time_steps), int(precision)),
                    "saveCheckPoint")

    def load_checkpoint(self, path):
        self._check(self.lib.load_checkpoint(self.h, _b(path)), "loadCheckPoint")

    def particle_stress(self):
        """Unscaled stress sums of the last step, (n, 3, 3), ascending id order."""
        n = self.lib.get_num_particles(self.h)
        out = np.zeros(9 * max(n, 1))
        self._check(self.lib.get_particle_stress(self.h, _dp(out)), "getParticleStress")
        return out[: 9 * n].reshape(n, 3, 3)

    def add_nrot_elastic_group(self, name, k, scaling=True):
        self._check(self.lib.add_nrot_elastic_group(self.h, _b(name), float(k), int(scaling)),
                    "createInteractionGroup(NRotElasticPrms)")

    def add_rot_elastic_group(self, name, kr, scaling=True):
        self._check(self.lib.add_rot_elastic_group(self.h, _b(name), float(kr), int(scaling)),
                    "createInteractionGroup(RotElasticPrms)")

    def add_rot_friction_group(self, name, prms):
        c = EspRotFrictionPrms()
        for f, _ in EspRotFrictionPrms._fields_:
            setattr(c, f, prms[f])
        self._check(self.lib.add_rot_friction_group(self.h, _b(name), ctypes.byref(c)),
                    "createInteractionGroup(RotFrictionPrms)")

    def add_exclusion(self, excluding, excluded):
        self._check(self.lib.add_exclusion(self.h, _b(excluding), _b(excluded)), "createExclusion")

    def add_wall(self, name, origin, normal):
        o = np.asarray(origin, np.float64)
        n = np.asarray(normal, np.float64)
        self._check(self.lib.add_wall(self.h, _b(name), _dp(o), _dp(n)), "createWall")

    def add_elastic_wall_group(self, name, wall, k):
        self._check(self.lib.add_elastic_wall_group(self.h, _b(name), _b(wall), float(k)),
                    "createInteractionGroup(NRotElasticWallPrms)")

    def set_interaction_group_tags(self, name, tag1, mask1, tag2, mask2):
        self._check(self.lib.set_interaction_group_tags(self.h, _b(name), int(tag1), int(mask1),
                                                        int(tag2), int(mask2)),
                    "createInteractionGroupTagged")

    def set_wall_group_tags(self, name, tag, mask):
        self._check(self.lib.set_wall_group_tags(self.h, _b(name), int(tag), int(mask)),
                    "createInteractionGroupTagged(NRotElasticWallPrms)")

    def add_bonded_wall_group(self, name, wall, k, tag, mask):
        self._check(self.lib.add_bonded_wall_group(self.h, _b(name), _b(wall), float(k), int(tag),
                                                   int(mask)),
                    "createInteractionGroup(NRotBondedWallPrms)")

    def add_gravity(self, name, g):
        g = np.asarray(g, np.float64)
        self._check(self.lib.add_gravity(self.h, _b(name), _dp(g)), "createInteractionGroup(GravityPrms)")

    def set_particle_vel(self, pid, v):
        v = np.asarray(v, np.float64)
        self._check(self.lib.set_particle_vel(self.h, int(pid), _dp(v)), "setParticleVelocity")

    def set_particle_ang_vel(self, pid, w):
        w = np.asarray(w, np.float64)
        self._check(self.lib.set_particle_ang_vel(self.h, int(pid), _dp(w)),
                    "setParticleAngularVelocity")

    def add_damping_group(self, kind, name, viscosity, max_iterations):
        """kind: 0 Damping, 1 RotDamping, 2 LocalDamping, 3 RotLocalDamping."""
        self._check(self.lib.add_damping_group(self.h, int(kind), _b(name), float(viscosity),
                                               int(max_iterations)),
                    "createInteractionGroup(DampingPrms)")

    def move_wall_by(self, name, d):
        d = np.asarray(d, np.float64)
        self._check(self.lib.move_wall_by(self.h, _b(name), _dp(d)), "moveWallBy")

    def check_state(self):
        """Raises if the device bookkeeping is inconsistent (esp_check_state)."""
        self._check(self.lib.check_state(self.h), "checkState")

    def apply_force_to_wall_noalloc(self, name, f):
        f = np.ascontiguousarray(f, np.float64)
        self._check(self.lib.apply_force_to_wall_noalloc(self.h, _b(name), _dp(f)),
                    "applyForceToWall (no allocation)")

    def apply_force_to_wall(self, name, f):
        f = np.ascontiguousarray(f, np.float64)
        self._check(self.lib.apply_force_to_wall(self.h, _b(name), _dp(f)), "applyForceToWall")

    def set_wall_step_motion(self, name, d):
        d = np.asarray(d, np.float64)
        self._check(self.lib.set_wall_step_motion(self.h, _b(name), _dp(d)), "setWallStepMotion")

    # ---- stepping ----
    def run(self, nsteps):
        self._check(self.lib.run_time_steps(self.h, int(nsteps)), "runTimeStep")

    # ---- queries ----
    def num_particles(self):
        return self._check(self.lib.get_num_particles(self.h), "getNumParticles")

    def time_step(self):
        t = ctypes.c_longlong(0)
        self._check(self.lib.get_time_step(self.h, ctypes.byref(t)), "getTimeStep")
        return t.value

    def stats(self):
        s = EspStats()
        self._check(self.lib.get_stats(self.h, ctypes.byref(s)), "getStats")
        return s.as_dict()

    def ids(self):
        out = np.zeros(self.num_particles(), np.int32)
        self._check(self.lib.get_particle_ids(self.h, _ip(out)), "getParticleIds")
        return out

    def field(self, field):
        if isinstance(field, str):
            field = FIELD_NAMES[field]
        n = self.num_particles()
        w = FIELD_WIDTH[field]
        out = np.zeros(n * w, np.float64)
        self._check(self.lib.get_particle_data(self.h, field, _dp(out)), "getParticleData")
        return out.reshape(n, w) if w > 1 else out

    def wall_position(self, name):
        out = np.zeros(3)
        self._check(self.lib.get_wall_position(self.h, _b(name), _dp(out)), "getWallPosition")
        return out

    def wall_force(self, name):
        out = np.zeros(3)
        self._check(self.lib.get_wall_force(self.h, _b(name), _dp(out)), "getWallForce")
        return out

    def wall_force_history(self, name, n):
        out = np.zeros(3 * n)
        m = self._check(self.lib.get_wall_force_history(self.h, _b(name), int(n), _dp(out)),
                        "getWallForceHistory")
        return out[: 3 * m].reshape(m, 3)

    def interaction_count(self, name):
        return self.lib.get_interaction_count(self.h, _b(name))

    def interactions(self, name):
        n = self.interaction_count(name)
        if n < 0:
            raise RuntimeError("unknown interaction group %r" % name)
        a = np.zeros(max(n, 1), np.int32)
        b = np.zeros(max(n, 1), np.int32)
        m = self._check(self.lib.get_interactions(self.h, _b(name), _ip(a), _ip(b)),
                        "getInteractions")
        return np.stack([a[:m], b[:m]], axis=1)

    def track_bond_criteria(self):
        self._check(self.lib.track_bond_criteria(self.h), "createFieldSaver(breaking_criterion)")

    def bond_criteria(self, name):
        """(ids (n, 2), criterion (n,)) of the live bonds in list order."""
        n = self._check(self.lib.get_bond_criteria(self.h, _b(name), None, None, None),
                        "getBondCriteria")
        a = np.zeros(max(n, 1), np.int32)
        b = np.zeros(max(n, 1), np.int32)
        c = np.zeros(max(n, 1))
        m = self._check(self.lib.get_bond_criteria(self.h, _b(name), _ip(a), _ip(b), _dp(c)),
                        "getBondCriteria")
        return np.stack([a[:m], b[:m]], 1), c[:m]

    def friction_state(self, name):
        n = self.interaction_count(name)
        ff = np.zeros(3 * max(n, 1))
        fl = np.zeros(max(n, 1), np.int32)
        m = self._check(self.lib.get_friction_state(self.h, _b(name), _dp(ff), _ip(fl)),
                        "getFrictionState")
        return ff[: 3 * m].reshape(m, 3), fl[:m]

    def neighbour_pairs(self):
        n = self._check(self.lib.get_neighbour_pairs(self.h, None, None), "getNeighbourPairs")
        a = np.zeros(max(n, 1), np.int32)
        b = np.zeros(max(n, 1), np.int32)
        self._check(self.lib.get_neighbour_pairs(self.h, _ip(a), _ip(b)), "getNeighbourPairs")
        return np.stack([a[:n], b[:n]], axis=1)

    def bond_breaks(self):
        n = self._check(self.lib.get_bond_breaks(self.h, None, None, None), "getBondBreaks")
        s = np.zeros(max(n, 1), np.int32)
        a = np.zeros(max(n, 1), np.int32)
        b = np.zeros(max(n, 1), np.int32)
        self._check(self.lib.get_bond_breaks(self.h, _ip(s), _ip(a), _ip(b)), "getBondBreaks")
        return np.stack([s[:n], a[:n], b[:n]], axis=1)

    def set_kernel_timing(self, enable):
        self._check(self.lib.set_kernel_timing(self.h, int(bool(enable))), "setKernelTiming")

    def kernel_times(self):
        """{class: (total_ms, launches)} since timing was enabled (resets)."""
        ms = np.zeros(len(KERNEL_CLASSES))
        n = np.zeros(len(KERNEL_CLASSES), np.int64)
        self._check(self.lib.get_kernel_times(
            self.h, _dp(ms), n.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))),
            "getKernelTimes")
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(KERNEL_CLASSES)}

    def rebuild_steps(self):
        n = self._check(self.lib.get_rebuild_steps(self.h, None), "getRebuildSteps")
        s = np.zeros(max(n, 1), np.int32)
        self._check(self.lib.get_rebuild_steps(self.h, _ip(s)), "getRebuildSteps")
        return s[:n]

    def track_pair_events(self, name):
        self._check(self.lib.track_pair_events(self.h, _b(name)), "trackPairEvents")

    def pair_events(self, name):
        """(step, id1, id2, kind) arrays of a tracked group's pair events, kind
        +1 for a created interaction and -1 for a removed one."""
        n = self._check(self.lib.get_pair_events(self.h, _b(name), None, None, None, None),
                        "getPairEvents")
        v = [np.zeros(max(n, 1), np.int32) for _ in range(4)]
        self._check(self.lib.get_pair_events(self.h, _b(name), *[_ip(a) for a in v]),
                    "getPairEvents")
        return tuple(a[:n] for a in v)

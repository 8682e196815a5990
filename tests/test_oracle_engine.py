"""Pins the oracle engine (neighbour table, storages, step order, integrator)
with brute force, closed forms and scripted semantics of the reference."""
import math

import numpy as np
import pytest


def make_sim(pkg, orc, ptype="NRotSphere", grid=1.2, alpha=0.2, dt=1e-3,
             lo=(0, 0, 0), hi=(10, 10, 10)):
    sim = orc.oracle_lsm_class(pkg)()
    sim.initNeighbourSearch(ptype, grid, alpha)
    sim.setTimeStepSize(dt)
    sim.setSpatialDomain(pkg.BoundingBox(pkg.Vec3(*lo), pkg.Vec3(*hi)))
    return sim


def brute_pairs(pos, rad, alpha):
    n = len(pos)
    out = []
    for i in range(n):
        for j in range(i + 1, n):
            d = pos[i] - pos[j]
            dist2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
            dmax = rad[i] + rad[j] + alpha
            if dist2 <= dmax * dmax:
                out.append((i, j))
    return sorted(out)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_full_list_equals_brute_force(pkg, orc, seed):
    """getFullList over the 27-cell stencil == all pairs within r_i + r_j + alpha."""
    rng = np.random.default_rng(seed)
    n = 400
    rad = rng.uniform(0.2, 0.5, n)
    pos = rng.uniform(0.5, 9.5, (n, 3))
    alpha = 0.2
    sim = make_sim(pkg, orc, grid=2 * 0.5 + alpha, alpha=alpha)
    sim.createParticleArrays(np.arange(n), pos, rad)
    sim.createInteractionGroup(pkg.NRotElasticPrms("e", 1.0))
    sim.runTimeStep()
    got = sorted(map(tuple, sim.engine.neighbour_pairs().tolist()))
    assert got == brute_pairs(pos, rad, alpha)
    # NE stores every candidate pair (no exclusions)
    inter = sorted(map(tuple, sim.getInteractionPairs("e").tolist()))
    assert inter == got


def test_particles_outside_grid_dropped(pkg, orc):
    sim = make_sim(pkg, orc)
    # the table covers one padding cell beyond the (grown) domain
    sim.createParticleArrays([0, 1, 2], [[5, 5, 5], [-0.5, 5, 5], [-3, 5, 5]], [0.3] * 3)
    assert sim.getNumParticles() == 2
    assert list(sim.getParticleIds()) == [0, 1]


def test_gravity_free_fall_closed_form(pkg, orc):
    """Symplectic Euler under constant g: v_n = n g dt, y_n = y0 + g dt^2 n(n+1)/2."""
    g, dt, n = -9.81, 1e-3, 300
    sim = make_sim(pkg, orc, dt=dt, hi=(10, 1000, 10), lo=(0, -1000, 0), grid=2.0)
    sim.createParticleArrays([7], [[5, 500, 5]], [0.4], mass=[2.0])
    sim.createInteractionGroup(pkg.GravityPrms("g", pkg.Vec3(0, g, 0)))
    sim.runTimeSteps(n)
    v = sim.getParticleData("vel")[0]
    x = sim.getParticleData("pos")[0]
    assert abs(v[1] - n * g * dt) < 1e-12 * n
    assert abs(x[1] - (500 + g * dt * dt * n * (n + 1) / 2)) < 1e-10
    np.testing.assert_allclose(sim.getParticleData("force")[0], [0, 2.0 * g, 0])


def test_free_rotation_closed_form(pkg, orc):
    """A moment-free sphere spinning at w: each step q <- normalise(q + dt/2 (0,w) q)
    rotates by 2 atan(|w| dt/2) about w."""
    dt, n = 1e-2, 250
    w = np.array([0.3, -1.2, 2.0])
    sim = make_sim(pkg, orc, ptype="RotSphere", dt=dt, grid=2.0)
    sim.createParticleArrays([0], [[5, 5, 5]], [0.5])
    # angular velocity through the particle-state hook of the oracle
    eng = sim.engine
    st = np.zeros(22)
    st[9] = 1.0
    st[6:9] = w
    st[19] = 1.0
    st[20] = 1.0
    olib = orc.load(pkg)
    for _ in range(n):
        olib.math_rot_particle_step(st.ctypes.data_as(__import__("ctypes").POINTER(
            __import__("ctypes").c_double)), dt)
    ang = n * 2 * math.atan(np.linalg.norm(w) * dt / 2)
    axis = w / np.linalg.norm(w)
    expect = np.array([math.cos(ang / 2), *(math.sin(ang / 2) * axis)])
    np.testing.assert_allclose(st[9:13], expect, atol=1e-12)
    np.testing.assert_allclose(st[6:9], w, rtol=0)
    assert eng.num_particles() == 1


def test_wall_contact_closed_form(pkg, orc):
    """CElasticWallInteraction: F = k pi r (r - d) n; the wall force is the sum
    over inner candidates."""
    sim = make_sim(pkg, orc, grid=2.0, lo=(-10, -10, -10), hi=(10, 10, 10))
    r, d, k = 0.5, 0.4, 3.0
    sim.createParticleArrays([0], [[1.0, d, 2.0]], [r])
    sim.createWall("w", pkg.Vec3(0, 0, 0), pkg.Vec3(0, 1, 0))
    sim.createInteractionGroup(pkg.NRotElasticWallPrms("wi", "w", k))
    sim.runTimeStep()
    f = k * 3.141592654 * r * (r - d)
    np.testing.assert_allclose(sim.getParticleData("force")[0], [0, f, 0], rtol=1e-14)
    np.testing.assert_allclose(list(sim.getWallForce("w")), [0, f, 0], rtol=1e-14)


def two_particle_bonded(pkg, orc, strength):
    sim = make_sim(pkg, orc, ptype="RotSphere", grid=2.4, alpha=0.2, dt=1e-3)
    sim.createParticleArrays([3, 9], [[5.0, 5, 5], [6.0, 5, 5]], [0.5, 0.5],
                             vel=[[-1.0, 0, 0], [1.0, 0, 0]])
    sim.createConnectionArrays([9], [3], [0])
    sim.createInteractionGroup(pkg.RotBondPrms("bonds", 1e3, 1e3, 1e3, 1e3, strength,
                                               strength, strength, strength, 0))
    sim.createInteractionGroup(pkg.RotFrictionPrms("friction", 1e3, 0.5, 0.5, 1e3))
    sim.createExclusion("bonds", "friction")
    return sim


def test_bond_break_event_and_flagged_pairs(pkg, orc):
    """A bond that exceeds its strength is removed at the next step's update
    (RotBondedInteraction::broken on the cached forces); its two particles are
    flagged, so the friction group takes the pair from the flagged list
    (no distance test) - and the bond no longer excludes it."""
    sim = two_particle_bonded(pkg, orc, strength=0.05)
    # step 1 starts undeformed (zero cached forces); step 2 stretches the bond;
    # step 3's update sees step 2's forces and removes it
    sim.runTimeSteps(2)
    assert sim.getNumInteractions("bonds") == 1
    assert sim.getNumInteractions("friction") == 0  # excluded while bonded
    sim.runTimeStep()
    br = sim.getBondBreaks()
    assert br.tolist() == [[3, 3, 9]]
    assert sim.getNumInteractions("bonds") == 0
    assert sim.getInteractionPairs("friction").tolist() == [[3, 9]]


def test_unbreakable_bond_holds(pkg, orc):
    sim = two_particle_bonded(pkg, orc, strength=1e9)
    sim.runTimeSteps(20)
    assert sim.getNumInteractions("bonds") == 1
    assert len(sim.getBondBreaks()) == 0


def test_friction_history_survives_rebuild_only_when_touching(pkg, orc):
    """PIS_ED::update on a rebuilt table keeps persistent (touching) contacts with
    their shear spring; a non-touching stored contact is re-created with none."""
    sim = make_sim(pkg, orc, ptype="RotSphere", grid=2.4, alpha=0.2, dt=1e-3)
    # pair A (0,1) touching and sliding; pair B (2,3) within alpha but apart;
    # particle 4 moves fast to force rebuilds
    sim.createParticleArrays(
        [0, 1, 2, 3, 4],
        [[2, 2, 2], [2.95, 2, 2], [2, 6, 2], [3.1, 6, 2], [7, 7, 7]],
        [0.5] * 5,
        vel=[[0, 0, 0], [0, 0.5, 0], [0, 0, 0], [0, 0, 0], [40.0, 0, 0]])
    sim.createInteractionGroup(pkg.RotFrictionPrms("f", 1e3, 0.5, 0.6, 1e3))
    sim.runTimeSteps(5)
    eng = sim.engine
    assert len(eng.rebuild_steps()) >= 2
    pairs = sim.getInteractionPairs("f").tolist()
    ff, fl = eng.friction_state("f")
    k = pairs.index([0, 1])
    assert fl[k] & 2 and np.linalg.norm(ff[k]) > 0  # history kept across rebuilds
    k = pairs.index([2, 3])
    assert fl[k] == 0 and np.all(ff[k] == 0)


def test_rebuild_trigger_half_verlet_distance(pkg, orc):
    """checkNeighbors rebuilds once some particle moved more than verletDist/2."""
    dt, alpha, v = 1e-3, 0.2, 9.0
    sim = make_sim(pkg, orc, grid=2.0, alpha=alpha, dt=dt, hi=(100, 10, 10))
    sim.createParticleArrays([0], [[5, 5, 5]], [0.5], vel=[[v, 0, 0]])
    sim.runTimeSteps(40)
    steps = sim.engine.rebuild_steps().tolist()
    # a rebuild at step s resets x_old before that step's motion; the check at
    # step s' sees (s' - s) v dt, so rebuilds repeat every floor(alpha/2/(v dt)) + 1 steps
    per = math.floor((alpha / 2) / (v * dt)) + 1
    expect = [1] + list(range(1 + per, 41, per))
    assert steps == expect

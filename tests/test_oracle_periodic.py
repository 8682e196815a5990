"""Pins the oracle's circular x boundary and the wall servo (applyForceToWall)
with brute force and closed forms.

Circular x (setSpatialDomain(bbox, circDimList=[True, False, False]),
TSubLattice::initNeighborTable(min, max, circ), SubLattice.hpp:495-564): whole
cells span the x range L, a particle that crossed an x end moves by -+L at the
next neighbour-table rebuild (CParticle::setCircular, Particle.cpp:336-342),
and pairs interact through the nearest image.

Servo (CEWallInteractionGroup::applyForce, EWallInteractionGroup.hpp:93-164;
CBWallInteractionGroup::applyForce, BWallInteractionGroup.hpp:71-145): the
wall moves along the force until the wall force balances it; with every
candidate in contact the first step solves the linear balance exactly."""
import math

import numpy as np
import pytest


def make_sim(pkg, orc, ptype="NRotSphere", grid=1.2, alpha=0.2, dt=1e-3, lo=(0, 0, 0),
             hi=(12, 6, 6), circ=True):
    sim = orc.oracle_lsm_class(pkg)()
    sim.initNeighbourSearch(ptype, grid, alpha)
    sim.setTimeStepSize(dt)
    sim.setSpatialDomain(pkg.BoundingBox(pkg.Vec3(*lo), pkg.Vec3(*hi)),
                         circDimList=[circ, False, False])
    return sim


def min_image_pairs(pos, rad, alpha, L):
    n = len(pos)
    out = []
    for i in range(n):
        for j in range(i + 1, n):
            d = pos[i] - pos[j]
            d[0] -= L * round(d[0] / L)
            if d @ d <= (rad[i] + rad[j] + alpha) ** 2:
                out.append((i, j))
    return sorted(out)


@pytest.mark.parametrize("seed", [1, 2])
def test_full_list_equals_minimum_image_brute_force(pkg, orc, seed):
    rng = np.random.default_rng(seed)
    n = 500
    L = 12.0
    rad = rng.uniform(0.2, 0.5, n)
    pos = np.column_stack([rng.uniform(0.0, L, n), rng.uniform(0.5, 5.5, n),
                           rng.uniform(0.5, 5.5, n)])
    sim = make_sim(pkg, orc, grid=1.2, alpha=0.2)
    sim.createParticleArrays(np.arange(n), pos, rad)
    sim.createInteractionGroup(pkg.NRotElasticPrms("e", 1.0))
    sim.runTimeStep()
    got = sorted(map(tuple, sim.engine.neighbour_pairs().tolist()))
    want = min_image_pairs(pos, rad, 0.2, L)
    assert got == want
    assert any(abs(pos[i, 0] - pos[j, 0]) > L / 2 for i, j in want)  # pairs across the edge


def test_cell_size_adjusted_to_whole_cells(pkg, orc):
    """gridSpacing 1.3 does not divide L = 12: the range becomes 12 / 9."""
    sim = make_sim(pkg, orc, grid=1.3)
    sim.createParticleArrays([0, 1], [[0.2, 3, 3], [12.0 - 0.2, 3, 3]], [0.3, 0.3])
    sim.createInteractionGroup(pkg.NRotElasticPrms("e", 1.0))
    sim.runTimeStep()
    # the two are 0.4 apart through the edge: a pair of the full list
    assert sim.engine.neighbour_pairs().tolist() == [[0, 1]]


def test_particle_wraps_at_rebuild(pkg, orc):
    """A particle moving in +x leaves through x = L and is moved by -L at the
    next rebuild; the free flight is otherwise unchanged."""
    v, dt = 10.0, 1e-3
    sim = make_sim(pkg, orc, dt=dt)
    sim.createParticleArrays([0], [[11.9, 3.0, 3.0]], [0.3], vel=[[v, 0.0, 0.0]])
    x = 11.9
    wrapped = False
    for _ in range(60):
        sim.runTimeStep()
        x = x + v * dt  # symplectic Euler without force: x += v dt
        px = sim.getParticleData("pos")[0, 0]
        if px < 6.0:
            wrapped = True
            assert px == x - 12.0 or abs(px - (x - 12.0)) < 1e-12
        else:
            assert abs(px - x) < 1e-12
    assert wrapped


def test_contact_across_edge_closed_form(pkg, orc):
    """Two spheres overlapping by delta through the x ends: the elastic
    repulsion k (r0/2)^2 / r0 delta pushes each away from the other across the
    edge (+x for the one near x = 0)."""
    r, delta, k, dt, m = 0.35, 0.1, 20.0, 1e-3, 1.0
    gap = 2 * r - delta
    sim = make_sim(pkg, orc, dt=dt)
    sim.createParticleArrays([0, 1], [[0.5 * gap - 0.1, 3, 3], [12.0 - 0.5 * gap - 0.1, 3, 3]],
                             [r, r], mass=[m, m])
    sim.createInteractionGroup(pkg.NRotElasticPrms("e", k))
    sim.runTimeStep()
    F = sim.getParticleData("vel") * (m / dt)
    want = k * (0.5 * 2 * r) ** 2 / (2 * r) * delta
    np.testing.assert_allclose(F[0], [want, 0, 0], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(F[1], [-want, 0, 0], rtol=1e-9, atol=1e-12)


def test_bond_across_edge(pkg, orc):
    """A bond created through the x ends has its initial vector to the image:
    no force at creation; stretched by delta it pulls with kr_eff delta."""
    r, dt, m, kr, delta = 0.5, 1e-3, 1.0, 100.0, 0.002
    for stretch in (0.0, delta):
        sim = make_sim(pkg, orc, ptype="RotSphere", dt=dt)
        sim.createParticleArrays([0, 1], [[0.3 + stretch, 3, 3], [12.0 - 0.7, 3, 3]], [r, r],
                                 mass=[m, m])
        sim.createConnectionArrays([0], [1], [0])
        sim.createInteractionGroup(pkg.RotBondPrms("b", kr, 10.0, 10.0, 10.0, 1e9, 1e9, 1e9, 1e9,
                                                   tag=0, scaling=False))
        sim.runTimeStep()
        F = sim.getParticleData("vel") * (m / dt)
        if stretch == 0.0:
            assert np.max(np.abs(F)) < 1e-9
        else:
            # particle 0 (x = 0.3 + delta) is pulled back across the edge: -x
            np.testing.assert_allclose(F[0], [-kr * delta, 0, 0], rtol=1e-6)
            np.testing.assert_allclose(F[1], [kr * delta, 0, 0], rtol=1e-6)


def test_momentum_conserved_through_the_edge(pkg, orc):
    """Pair forces across the x ends obey Newton's third law: a closed periodic
    box (no walls, no gravity) keeps its total momentum."""
    rng = np.random.default_rng(4)
    n = 300
    rad = rng.uniform(0.3, 0.5, n)
    pos = np.column_stack([rng.uniform(0, 12, n), rng.uniform(0.6, 5.4, n),
                           rng.uniform(0.6, 5.4, n)])
    vel = rng.normal(0, 1.0, (n, 3))
    mass = rad ** 3
    sim = make_sim(pkg, orc, ptype="RotSphere", dt=1e-3)
    sim.createParticleArrays(np.arange(n), pos, rad, mass=mass, vel=vel)
    sim.createInteractionGroup(pkg.RotFrictionPrms("f", 50.0, 0.5, 0.5, 20.0))
    p0 = (mass[:, None] * vel).sum(0)
    sim.runTimeSteps(300)
    ids = sim.getParticleIds()
    assert len(ids) == n  # nobody lost through the circular ends
    p1 = (sim.getParticleData("mass")[:, None] * sim.getParticleData("vel")).sum(0)
    assert np.max(np.abs(p1[1:] - p0[1:])) < 1e-9 * np.abs(mass[:, None] * vel).sum()
    assert abs(p1[0] - p0[0]) < 1e-9 * np.abs(mass[:, None] * vel).sum()


# ---------------------------------------------------------------------------
# wall servo
# ---------------------------------------------------------------------------
def _layer(pkg, orc, n=12, seed=3, lo=(0, 0, 0), hi=(8, 6, 8), circ=False):
    rng = np.random.default_rng(seed)
    rad = rng.uniform(0.3, 0.5, n)
    xs = np.linspace(0.8, 7.2, n)
    pos = np.column_stack([xs, 2.0 + rng.uniform(-0.05, 0.05, n), rng.uniform(1.0, 7.0, n)])
    sim = make_sim(pkg, orc, ptype="NRotSphere", grid=1.2, lo=lo, hi=hi, circ=circ)
    sim.createParticleArrays(np.arange(n), pos, rad, tag=np.ones(n, np.int32))
    sim.setParticleNonDynamic(1)
    return sim, pos, rad


def test_servo_elastic_wall_linear_balance(pkg, orc):
    """All candidates touch the wall: one step of applyForce solves
    sum k_i (r_i - dist_i') = |F| exactly, so the wall ends at
    y - (|F| - sum k_i (r_i - dist_i)) / sum k_i."""
    sim, pos, rad = _layer(pkg, orc)
    top = float(np.min(pos[:, 1] + rad)) - 0.02  # every particle overlaps the wall
    sim.createWall("top", pkg.Vec3(0, top, 0), pkg.Vec3(0, -1, 0))
    k = 5.0
    sim.createInteractionGroup(pkg.NRotElasticWallPrms("tw", "top", k))
    sim.runTimeStep()  # wall candidates are listed at the first step's update
    keff = 3.141592654 * rad * k  # the literal of EWallInteraction.hpp:33
    dist = top - pos[:, 1]
    assert np.all(dist < rad)
    Fm = 3.0
    d = (Fm - np.sum(keff * (rad - dist))) / np.sum(keff)
    sim.applyForceToWall("tw", pkg.Vec3(0, -Fm, 0))
    y = sim.getWallPosition("top")[1]
    assert abs(y - (top - d)) < 1e-12 * max(1.0, abs(top))


def test_servo_elastic_wall_without_contact(pkg, orc):
    """No candidate touches: the wall advances 1e-5 per iteration for the
    50 iterations the loop allows."""
    sim, pos, rad = _layer(pkg, orc)
    top = float(np.max(pos[:, 1] + rad)) + 0.05
    sim.createWall("top", pkg.Vec3(0, top, 0), pkg.Vec3(0, -1, 0))
    sim.createInteractionGroup(pkg.NRotElasticWallPrms("tw", "top", 5.0))
    sim.runTimeStep()
    sim.applyForceToWall("tw", pkg.Vec3(0, -1.0, 0))
    y = top
    for _ in range(50):
        y = y + (-1.0) * 1e-5
    assert sim.getWallPosition("top")[1] == y


def test_servo_bonded_wall_balance(pkg, orc):
    """A bonded wall group whose candidates are all bonded: the wall force is
    linear in the wall's displacement, so after applyForce the bonded forces
    along the normal balance the applied force to within K 1e-5."""
    sim, pos, rad = _layer(pkg, orc)
    top = float(np.max(pos[:, 1] + rad)) + 0.1
    sim.createWall("top", pkg.Vec3(0, top, 0), pkg.Vec3(0, -1, 0))
    k = 5.0
    sim.createInteractionGroup(pkg.NRotBondedWallPrms("tb", "top", k, 1))
    sim.runTimeStep()
    Fm = 2.0
    sim.applyForceToWall("tb", pkg.Vec3(0, -Fm, 0))
    dw = sim.getWallPosition("top")[1] - top  # wall displacement (particles fixed)
    kb = k * 3.1415926536 * rad
    # bonded force on the wall: -(particle disp - wall disp) k = dw k per particle
    fwall = np.sum(kb * dw)
    assert abs(fwall - (-Fm)) <= 1e-5 * np.sum(kb) + 1e-12
    assert dw < 0

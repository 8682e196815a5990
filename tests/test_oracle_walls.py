"""Pins the oracle's bonded wall groups (CBWallInteractionGroup,
BWallInteractionGroup.hpp; CBondedWallInteraction::calcForces,
BWallInteraction.hpp:22-52) with closed forms: a candidate particle whose tag
matches is pulled by -(its displacement since creation - the wall's
displacement) k', k' = k * 3.1415926536 r in 3-D and k in 2-D, and the wall
reports the sum of those forces; candidates whose tag does not match interact
elastically instead (no force while they do not touch the wall)."""
import math

import numpy as np
import pytest


def model(pkg, orc, tags, two_d=False, mask=-1, wall_step=(0.2e-3, 0.0, 0.0), held=True):
    sim = orc.oracle_lsm_class(pkg)()
    sim.initNeighbourSearch("NRotSphere", 2.0, 0.2)
    sim.setTimeStepSize(1e-3)
    if two_d:
        sim.force2dComputations(True)
    sim.setSpatialDomain(pkg.BoundingBox(pkg.Vec3(-10, -10, -10), pkg.Vec3(10, 10, 10)))
    n = len(tags)
    pos = [[3.0 * i, 0.8, 0.0] for i in range(n)]
    sim.createParticleArrays(range(n), pos, [0.5] * n, mass=[1.0] * n, tag=tags)
    if held:
        sim.setParticleNonDynamic(7)  # tag 7: held in place
    sim.createWall("floor", pkg.Vec3(0, 0, 0), pkg.Vec3(0, 1, 0))
    sim.createInteractionGroup(pkg.NRotBondedWallPrms("bw", "floor", 40.0, 7, mask))
    sim.setWallStepMotion("floor", pkg.Vec3(*wall_step))
    return sim


@pytest.mark.parametrize("two_d", [False, True])
def test_held_particle_follows_wall_displacement(pkg, orc, two_d):
    # particle 0 (tag 7) is bonded and held; particle 1 (tag 2) is elastic and
    # above the wall (0.8 > r): no force
    sim = model(pkg, orc, [7, 2], two_d=two_d)
    steps = 25
    sim.runTimeSteps(steps)
    d = steps * 0.2e-3
    kp = 40.0 * (1.0 if two_d else 3.1415926536 * 0.5)
    f = sim.getParticleData("force")
    np.testing.assert_allclose(f[0], [d * kp, 0.0, 0.0], rtol=1e-12)
    np.testing.assert_array_equal(f[1], [0.0, 0.0, 0.0])
    np.testing.assert_allclose(list(sim.getWallForce("floor")), [d * kp, 0.0, 0.0], rtol=1e-12)


def test_free_particle_relative_displacement(pkg, orc):
    # a bonded free particle moves too: the force is -(dx_p - dx_w) k'
    sim = model(pkg, orc, [7], wall_step=(0.0, 0.0, 0.0), held=False)
    sim.setParticleVelocity(0, pkg.Vec3(0.0, 0.0, 0.1))
    sim.runTimeStep()
    x1 = sim.getParticleData("pos")[0]
    sim.runTimeStep()  # the force of step 2 uses the displacement after step 1
    kp = 40.0 * 3.1415926536 * 0.5
    f = sim.getParticleData("force")[0]
    np.testing.assert_allclose(f, -(x1 - np.array([0.0, 0.8, 0.0])) * kp, rtol=1e-12)
    assert f[2] < 0


def test_tag_mask(pkg, orc):
    # mask 1: tags 7 and 3 match tag 7 & 1; tag 2 does not
    sim = model(pkg, orc, [7, 3, 2], mask=1)
    sim.setParticleNonDynamic(3)
    sim.runTimeSteps(10)
    f = sim.getParticleData("force")
    kp = 40.0 * 3.1415926536 * 0.5
    np.testing.assert_allclose(f[0], [10 * 0.2e-3 * kp, 0, 0], rtol=1e-12)
    np.testing.assert_allclose(f[1], [10 * 0.2e-3 * kp, 0, 0], rtol=1e-12)
    np.testing.assert_array_equal(f[2], [0.0, 0.0, 0.0])


def test_tagged_elastic_wall(pkg, orc):
    """createInteractionGroupTagged(NRotElasticWallPrms, tag, mask): only
    candidates with the tag feel the wall (TaggedEWallInteractionGroup.hpp:28-52);
    the force of the elastic law is k 3.141592654 r (r - dist) along the normal
    (EWallInteraction.hpp:26-67)."""
    sim = orc.oracle_lsm_class(pkg)()
    sim.initNeighbourSearch("NRotSphere", 2.0, 0.2)
    sim.setTimeStepSize(1e-4)
    sim.setSpatialDomain(pkg.BoundingBox(pkg.Vec3(-10, -10, -10), pkg.Vec3(10, 10, 10)))
    sim.createParticleArrays([0, 1, 2], [[0.0, 0.4, 0.0], [3.0, 0.4, 0.0], [6.0, 0.4, 0.0]],
                             [0.5] * 3, mass=[1.0] * 3, tag=[5, 4, 13])
    sim.createWall("floor", pkg.Vec3(0, 0, 0), pkg.Vec3(0, 1, 0))
    sim.createInteractionGroupTagged(pkg.NRotElasticWallPrms("w", "floor", 20.0), 5, 7)
    sim.runTimeStep()
    f = sim.getParticleData("force")
    push = 20.0 * 3.141592654 * 0.5 * (0.5 - 0.4)
    for k, hit in enumerate([True, False, True]):  # 5 & 7 == 5, 4 & 7 != 5, 13 & 7 == 5
        np.testing.assert_allclose(f[k], [0.0, push if hit else 0.0, 0.0], rtol=1e-12)
    np.testing.assert_allclose(list(sim.getWallForce("floor")), [0.0, 2 * push, 0.0], rtol=1e-12)

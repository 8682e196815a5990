"""Pins the oracle's damping groups (TSubLattice::calcForces applies them last,
SubLattice.hpp:2143-2147) with closed forms on single particles.

CDamping (Damping.hpp:87-107) iterates v_rel = v + s dt (f - v_rel visc m)
from zero while |change|^4 > 1e-12, at most maxIterations times, and a run
that reaches maxIterations applies nothing; the force is -visc m v_rel.  So:
- one iteration that already meets the limit (small v, f = 0) gives
  v_rel = v exactly and one step gives v (1 - visc dt);
- the same case with maxIterations = 1 applies nothing;
- a converged run approaches the fixed point v_rel* = (v + dt f / m) / (1 + a),
  a = visc dt, with |v_rel - v_rel*| <= a / (1 - a) * 1e-3 (the last change is
  at most 1e-3); f includes every other force of the step (here gravity).
CRotDamping (RotDamping.hpp:58-84) is the same on angular velocity with 1/I
and I (I = 0.4 m r^2).  CLocalDamping (LocalDamping.hpp:82-117) adds
-visc |f_c| sign(v_c) per component (nothing where v_c = 0)."""
import numpy as np
import pytest


def one(pkg, orc, ptype="RotSphere", dt=1e-3, vel=(0.0, 0.0, 0.0), angvel=None, mass=None,
        rad=0.5, groups=(), gravity=None, steps=1):
    sim = orc.oracle_lsm_class(pkg)()
    sim.initNeighbourSearch(ptype, 2.5, 0.2)
    sim.setTimeStepSize(dt)
    sim.setSpatialDomain(pkg.BoundingBox(pkg.Vec3(-20, -20, -20), pkg.Vec3(20, 20, 20)))
    sim.createParticleArrays([7], [[0.0, 0.0, 0.0]], [rad], mass=None if mass is None else [mass],
                             vel=[list(vel)])
    if angvel is not None:
        sim.setParticleAngularVelocity(7, pkg.Vec3(angvel))
    if gravity is not None:
        sim.createInteractionGroup(pkg.GravityPrms("g", pkg.Vec3(gravity)))
    for g in groups:
        sim.createInteractionGroup(g)
    sim.runTimeSteps(steps)
    return sim


def vel_of(sim):
    return sim.getParticleData("vel")[0]


def test_single_iteration_exact_and_max_iterations_one(pkg, orc):
    v0 = np.array([3e-4, -2e-4, 1e-4])  # |v|^4 < 1e-12: the first iteration meets the limit
    visc, dt = 2.0, 1e-3
    sim = one(pkg, orc, vel=v0, dt=dt, groups=[pkg.LinDampingPrms("d", visc, 2)])
    np.testing.assert_allclose(vel_of(sim), v0 * (1.0 - visc * dt), rtol=1e-14, atol=0)
    # the same run with maxIterations=1 reaches the limit: no damping at all
    sim1 = one(pkg, orc, vel=v0, dt=dt, groups=[pkg.LinDampingPrms("d", visc, 1)])
    np.testing.assert_array_equal(vel_of(sim1), v0)


@pytest.mark.parametrize("mass", [1.0, 8.0])
def test_converged_fixed_point_with_gravity(pkg, orc, mass):
    v0 = np.array([1.0, -0.5, 0.25])
    g = np.array([0.0, -200.0, 0.0])
    visc, dt = 500.0, 1e-3  # a = 0.5: several iterations
    a = visc * dt
    sim = one(pkg, orc, vel=v0, dt=dt, mass=mass, gravity=g,
              groups=[pkg.DampingPrms("Damping", "d", visc, 100)])
    vrel_star = (v0 + dt * g) / (1.0 + a)
    expect = v0 + dt * g - dt * visc * vrel_star
    bound = dt * visc * a / (1.0 - a) * 1e-3
    assert np.max(np.abs(vel_of(sim) - expect)) <= bound
    # gravity left out of the damped force would miss by dt g a / (1 + a)
    assert np.max(np.abs(dt * g * a / (1 + a))) > 10 * bound


def test_rot_damping_fixed_point(pkg, orc):
    w0 = np.array([2.0, -1.0, 0.5])
    visc, dt, m, r = 400.0, 1e-3, 3.0, 0.5  # I = 0.3, far from m
    a = visc * dt
    sim = one(pkg, orc, angvel=w0, dt=dt, mass=m, rad=r,
              groups=[pkg.RotDampingPrms("rd", visc, 100)])
    w1 = sim.getParticleData("angvel")[0]
    expect = w0 - dt * visc * w0 / (1.0 + a)
    assert np.max(np.abs(w1 - expect)) <= dt * visc * a / (1.0 - a) * 1e-3
    # the velocity is untouched by rotational damping
    np.testing.assert_array_equal(vel_of(sim), [0.0, 0.0, 0.0])


def test_local_damping_against_motion(pkg, orc):
    g = np.array([0.0, -10.0, 0.0])
    visc, dt, m = 0.3, 1e-3, 2.0
    v0 = np.array([0.5, -1.0, 0.0])  # x: force 0; y: against gravity's pull; z: v = 0
    sim = one(pkg, orc, vel=v0, dt=dt, mass=m, gravity=g,
              groups=[pkg.LocalDampingPrms("ld", visc)])
    f = sim.getParticleData("force")[0]
    # y: v < 0, f = m g < 0 -> damping -visc * (-|f|) = +visc m |g|
    np.testing.assert_allclose(f, [0.0, m * g[1] * (1.0 - visc), 0.0], rtol=1e-15, atol=0)


def test_damping_map_keys(pkg, orc):
    v0 = np.array([3e-4, 0.0, 0.0])
    dt = 1e-3
    # a second "Damping" group is dropped by the reference's map insert
    two = one(pkg, orc, vel=v0, dt=dt,
              groups=[pkg.LinDampingPrms("a", 2.0, 5), pkg.LinDampingPrms("b", 7.0, 5)])
    first = one(pkg, orc, vel=v0, dt=dt, groups=[pkg.LinDampingPrms("a", 2.0, 5)])
    np.testing.assert_array_equal(vel_of(two), vel_of(first))
    # RotDamping needs rotational particles
    with pytest.raises(RuntimeError):
        one(pkg, orc, ptype="NRotSphere", groups=[pkg.RotDampingPrms("r", 1.0)])

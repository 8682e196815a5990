"""Pins the oracle's particle stress sums (CParticle::applyForce,
Particle.cpp:391-400) and the BrittleBeamSC criterion
(BrittleBeamSCInteraction::broken, BrittleBeamSC.cpp:103-148) with closed forms:
a force f applied at point c changes a particle's stress sum by -f (x) (c - x);
an axially loaded bond therefore gives both particles sigma = -F a (n (x) n),
a_i = r_i |x2 - x1| / (r1 + r2), so in the bond frame sigma_n = -F a / V and
tau = 0 (V = 4/3 pi r^3)."""
import math

import numpy as np
import pytest


def make_sim(pkg, orc, ptype, dt, grid=2.5, alpha=0.2):
    sim = orc.oracle_lsm_class(pkg)()
    sim.initNeighbourSearch(ptype, grid, alpha)
    sim.setTimeStepSize(dt)
    sim.setSpatialDomain(pkg.BoundingBox(pkg.Vec3(-6, -6, -6), pkg.Vec3(6, 6, 6)))
    return sim


def test_contact_stress_closed_form(pkg, orc):
    r1, r2, d, k = 0.8, 0.5, 1.2, 7.0
    sim = make_sim(pkg, orc, "NRotSphere", 1e-4)
    sim.createParticleArrays([0, 1], [[0, 0, 0], [d, 0, 0]], [r1, r2])
    sim.createInteractionGroup(pkg.NRotElasticPrms("e", k, False))
    sim.runTimeStep()
    s = sim.getParticleStress()
    eq = r1 + r2
    f = k * (eq - d)  # repulsion along x
    e1 = np.zeros((3, 3))
    e2 = np.zeros((3, 3))
    e1[0, 0] = f * d * r1 / eq  # -(-f) * (c - x1), c - x1 = d r1 / eq
    e2[0, 0] = f * d * r2 / eq  # -(+f) * (c - x2), c - x2 = -d r2 / eq
    np.testing.assert_allclose(s[0], e1, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(s[1], e2, rtol=1e-14, atol=1e-14)


def axial_run(pkg, orc, n_hat, steps, sc=None, E=1e4, nu=0.25):
    """Two bonded particles, the second moving away along the bond axis."""
    r1, r2, v, dt = 0.6, 0.4, 0.5, 1e-4
    n_hat = np.asarray(n_hat, float)
    sim = make_sim(pkg, orc, "RotSphere", dt)
    x2 = n_hat * (r1 + r2)
    sim.createParticleArrays([0, 1], [[0.0, 0.0, 0.0], x2], [r1, r2], vel=[[0, 0, 0], v * n_hat])
    sim.createConnectionArrays([0], [1], [0])
    coh, ang, cut = sc if sc else (1e30, 30.0, 1e30)
    sim.createInteractionGroup(pkg.BrittleBeamSCPrms("b", E, nu, coh, ang, cut, 0))
    hist = []
    for s in range(steps):
        pos = sim.getParticleData("pos")
        sim.runTimeStep()
        f2 = sim.getParticleData("force")[1]
        hist.append((pos, f2, sim.getParticleStress()))
        if len(sim.getBondBreaks()):
            break
    return sim, hist, (r1, r2)


def predicted_sigma_n(pos, f2, radii):
    r1, r2 = radii
    d = pos[1] - pos[0]
    dist = math.sqrt(d @ d)
    n = d / dist
    F = -(f2 @ n)  # tension pulls particle 2 back
    vols = [4.0 / 3.0 * math.pi * r ** 3 for r in radii]
    a = [r * dist / (r1 + r2) for r in radii]
    return 0.5 * (-F * a[0] / vols[0] - F * a[1] / vols[1]), n, F, a


@pytest.mark.parametrize("n_hat", [(1.0, 0.0, 0.0), (1 / 3, 2 / 3, 2 / 3)])
def test_axial_bond_stress_closed_form(pkg, orc, n_hat):
    _, hist, radii = axial_run(pkg, orc, n_hat, 20)
    for pos, f2, s in hist[1:]:
        sn, n, F, a = predicted_sigma_n(pos, f2, radii)
        assert F > 0
        scale = F * max(a)
        np.testing.assert_allclose(s[0], -F * a[0] * np.outer(n, n), atol=1e-9 * scale)
        np.testing.assert_allclose(s[1], -F * a[1] * np.outer(n, n), atol=1e-9 * scale)


@pytest.mark.parametrize("n_hat", [(1.0, 0.0, 0.0), (1 / 3, 2 / 3, 2 / 3)])
@pytest.mark.parametrize("branch", ["tension", "mohr_coulomb"])
def test_sc_break_step(pkg, orc, n_hat, branch):
    """The bond breaks in the update after the first step whose mean normal
    stress passes the tension cutoff (sigma_n < -cutoff) or, with tau = 0, the
    Mohr-Coulomb line (0 > c + sigma_n tan(phi))."""
    _, hist, radii = axial_run(pkg, orc, n_hat, 30)
    sig = [predicted_sigma_n(p, f, radii)[0] for p, f, _ in hist]
    k = 12  # threshold between steps k and k+1 (1-based step numbers)
    assert sig[k] < sig[k - 1] < 0
    mid = 0.5 * (sig[k - 1] + sig[k])
    angle = 30.0
    if branch == "tension":
        sc = (1e30, angle, -mid)
    else:
        sc = (-mid * math.tan(angle / 57.2957795), angle, 1e30)
    sim, _, _ = axial_run(pkg, orc, n_hat, 30, sc=sc)
    ev = sim.getBondBreaks()
    assert ev.tolist() == [[k + 2, 0, 1]]

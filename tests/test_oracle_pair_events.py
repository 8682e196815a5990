"""The oracle's pair-event log against its own interaction lists.

An event is a change of a dynamic group's stored pairs from one step to the
next: pinned here by stepping the oracle one step at a time, reading the
group's interaction list after every step (the list parity tests pin against
the HIP engine) and taking the multiset difference of consecutive lists.
Cases: a bonded block under moving walls (bond breaks create friction pairs
between table rebuilds, rebuilds drop and add pairs), a block falling out of
the grid (erased particles take their pairs), a settling cloud with an
NRotElastic group (whole lists rebuilt)."""
import collections
import importlib

import numpy as np
import pytest


@pytest.fixture(scope="module")
def spm(pkg):
    return importlib.import_module("esys_particle_amd.specimens")


def expected_events(sim, name, steps):
    """Step the simulation one step at a time; events from the multiset
    difference of consecutive interaction lists."""
    prev = collections.Counter()
    ev = []
    for _ in range(steps):
        sim.runTimeSteps(1)
        t = sim.getTimeStep()
        pairs = sim.getInteractionPairs(name)
        cur = collections.Counter(map(tuple, pairs.tolist()))
        for k in set(prev) | set(cur):
            d = cur[k] - prev[k]
            ev += [(t, k[0], k[1], 1 if d > 0 else -1)] * abs(d)
        prev = cur
    return sorted(ev), prev


def check(sim, name, steps):
    exp, last = expected_events(sim, name, steps)
    got = list(zip(*[a.tolist() for a in sim.getPairEvents(name)]))
    assert got == exp
    # the running sum of the kinds is the number of stored interactions
    assert sum(e[3] for e in got) == sum(last.values()) == sim.getNumInteractions(name)
    return got


def oracle_sim(pkg, orc, spm, spec, names):
    o = orc.oracle_lsm_class(pkg)()
    spm.apply(o, spec)
    for n in names:
        o.trackPairEvents(n)
    return o


def lattice_spec(spm):
    spec = spm.lattice_block(nx=5, ny=8, nz=5, dt=1e-4, vel_sd=0.05, wall_speed=5.0)
    top = float(spec["pos"][:, 1].max() + spec["rad"].max()) - 0.01
    bot = float(spec["pos"][:, 1].min() - spec["rad"].max()) + 0.01
    w = spec["walls"]
    spec["walls"] = [(w[0][0], (0.0, bot, 0.0)) + w[0][2:], (w[1][0], (0.0, top, 0.0)) + w[1][2:]]
    return spec


def test_events_of_a_breaking_block(pkg, orc, spm):
    o = oracle_sim(pkg, orc, spm, lattice_spec(spm), ["friction"])
    got = check(o, "friction", 200)
    steps = {e[0] for e in got}
    rebuilds = set(o.engine.rebuild_steps().tolist())
    breaks = {int(s) for s in o.getBondBreaks()[0]} if len(o.getBondBreaks()) else set()
    assert any(e[3] < 0 for e in got) and any(e[3] > 0 for e in got)
    # pairs are created between rebuilds too: flagged pairs after bond breaks
    assert steps - rebuilds
    assert len(o.getBondBreaks()) > 50 and breaks


def test_events_of_erased_pairs(pkg, orc, spm):
    spec = spm.lattice_block(nx=5, ny=8, nz=5, dt=1e-4, vel_sd=0.05)
    spec["vel"] = spec["vel"] + np.array([0.0, -40.0, 0.0])
    spec["walls"] = []
    spec["gravity"] = (0.0, -9.81, 0.0)
    spec["domain"] = ((-1.0, 2.4, -1.0), spec["domain"][1])
    o = oracle_sim(pkg, orc, spm, spec, ["friction"])
    n0 = len(spec["ids"])
    got = check(o, "friction", 300)
    assert o.getNumParticles() < n0
    gone = set(spec["ids"].tolist()) - set(o.getParticleIds().tolist())
    assert any(e[3] < 0 and (e[1] in gone or e[2] in gone) for e in got)


def test_events_of_an_elastic_group(pkg, orc, spm):
    spec = spm.gravity_settling(n=300, seed=5, box=(6.0, 12.0, 6.0), dt=2e-3)
    spec["vel"] = np.random.default_rng(5).normal(0.0, 2.0, (300, 3))
    o = oracle_sim(pkg, orc, spm, spec, ["elastic"])
    got = check(o, "elastic", 150)
    assert len(o.engine.rebuild_steps()) > 2 and any(e[3] < 0 for e in got)


def test_tracking_rules(pkg, orc, spm):
    o = orc.oracle_lsm_class(pkg)()
    spm.apply(o, lattice_spec(spm))
    with pytest.raises(Exception):
        o.trackPairEvents("no_such_group")
    with pytest.raises(Exception):
        o.getPairEvents("friction")  # not tracked
    o.runTimeSteps(1)
    with pytest.raises(Exception):
        o.trackPairEvents("friction")  # after the first step


def erase_break_sim(cls, lsm, vy):
    """Four rotational spheres bonded to particle 0; the bond 0-1 starts 0.1
    beyond its equilibrium length with a normal strength of 1e-3, so its break
    criterion holds after step 1.  Particle 1 starts in the grid's bottom
    padding cell with velocity (0, vy, 0)."""
    s = cls()
    s.initNeighbourSearch("RotSphere", 1.2, 0.2)
    s.setTimeStepSize(1e-3)
    s.setSpatialDomain(lsm.BoundingBox(lsm.Vec3(-1, 2.4, -1), lsm.Vec3(6, 8, 6)))
    ids = np.array([0, 1, 2, 3], np.int32)
    pos = np.array([[2.0, 2.7, 2.0], [2.0, 1.6, 2.0], [3.0, 2.7, 2.0], [2.0, 3.7, 2.0]])
    rad = np.full(4, 0.5)
    vel = np.zeros((4, 3))
    vel[1] = [0.0, vy, 0.0]
    s.createParticleArrays(ids, pos, rad, mass=rad ** 3, tag=np.zeros(4, np.int32), vel=vel)
    s.createConnectionArrays(np.array([0, 0, 0]), np.array([1, 2, 3]), np.zeros(3, np.int32))
    st = np.array([[1e-3, 1e9, 1e9, 1e9], [1e9] * 4, [1e9] * 4])
    s.createInteractionGroup(lsm.RotBondPrms("bonded", 1e5, 3e4, 3e4, 3e4, 1, 1, 1, 1, 0,
                                             perBondStrengths=st))
    s.createInteractionGroup(lsm.RotFrictionPrms("friction", 1e5, 0.6, 0.6, 3e4))
    s.createExclusion("bonded", "friction")
    return s


def test_search_erases_before_the_break_test(pkg, orc):
    """CLatticeMaster::runOneStep searches neighbours before it updates the
    interactions (LatticeMaster.cpp:2602-2725): a particle that left the grid
    is erased with its bond (ntable.hpp:281-286, pi_storage_e.hpp:111-155)
    before PIS_EB::update tests the bond, so the bond whose criterion held
    gives no break event; the same bond breaks at step 2 when the particle
    stays in the grid."""
    cls = orc.oracle_lsm_class(pkg)
    stay = erase_break_sim(cls, pkg.lsm, 0.0)
    stay.runTimeSteps(3)
    assert stay.getBondBreaks().tolist() == [[2, 0, 1]]
    gone = erase_break_sim(cls, pkg.lsm, -1000.0)
    gone.runTimeSteps(3)
    assert gone.getParticleIds().tolist() == [0, 2, 3]
    assert gone.engine.rebuild_steps().tolist() == [1, 2]
    assert gone.getBondBreaks().tolist() == []
    assert gone.getInteractionPairs("bonded").tolist() == [[0, 2], [0, 3]]

"""Pair-event log, several bonded groups and the checkpoint's neighbour search
under particle erasure: the HIP engine against the oracle.

Bars: events, bond breaks, rebuild steps and interaction lists bit-exact;
trajectories within 1e-9 after 200-300 steps; tracking changes nothing
(positions bitwise equal to an untracked run)."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TRAJ_TOL = 1e-9


@pytest.fixture(scope="module")
def spm(pkg):
    return importlib.import_module("esys_particle_amd.specimens")


def breaking_block(spm):
    spec = spm.lattice_block(nx=6, ny=10, nz=6, dt=1e-4, vel_sd=0.05, wall_speed=5.0)
    top = float(spec["pos"][:, 1].max() + spec["rad"].max()) - 0.01
    bot = float(spec["pos"][:, 1].min() - spec["rad"].max()) + 0.01
    w = spec["walls"]
    spec["walls"] = [(w[0][0], (0.0, bot, 0.0)) + w[0][2:], (w[1][0], (0.0, top, 0.0)) + w[1][2:]]
    return spec


def falling_block(spm):
    """A bonded block falling out of the neighbour table's grid: erasures at
    rebuilds, bond breaks and flagged friction pairs between them."""
    spec = spm.lattice_block(nx=6, ny=10, nz=6, dt=1e-4, vel_sd=0.05)
    spec["vel"] = spec["vel"] + np.array([0.0, -20.0, 0.0])
    spec["walls"] = []
    spec["gravity"] = (0.0, -9.81, 0.0)
    spec["domain"] = ((-1.0, 2.4, -1.0), spec["domain"][1])
    return spec


def settling_cloud(spm):
    spec = spm.gravity_settling(n=600, seed=5, box=(8.0, 16.0, 8.0), dt=2e-3)
    spec["vel"] = np.random.default_rng(5).normal(0.0, 2.0, (600, 3))
    return spec


def tracked_pair(pkg, orc, spm, spec, names, steps, batch=32):
    g = pkg.LsmMpi(stepsPerBatch=batch)
    spm.apply(g, spec)
    o = orc.oracle_lsm_class(pkg)()
    spm.apply(o, spec)
    for n in names:
        g.trackPairEvents(n)
        o.trackPairEvents(n)
    print("case: n=%d steps=%d groups=%s" % (len(spec["ids"]), steps, names), flush=True)
    g.runTimeSteps(steps)
    o.runTimeSteps(steps)
    return g, o


def assert_same_events(g, o, name):
    eg, eo = g.getPairEvents(name), o.getPairEvents(name)
    for a, b in zip(eg, eo):
        np.testing.assert_array_equal(a, b)
    assert len(eo[0]) > 0
    assert int(eg[3].sum()) == g.getNumInteractions(name)
    np.testing.assert_array_equal(g.getInteractionPairs(name), o.getInteractionPairs(name))
    assert g.engine.rebuild_steps().tolist() == o.engine.rebuild_steps().tolist()
    np.testing.assert_array_equal(g.getBondBreaks(), o.getBondBreaks())


def test_events_of_a_breaking_block(pkg, orc, spm):
    spec = breaking_block(spm)
    g, o = tracked_pair(pkg, orc, spm, spec, ["friction"], 200)
    assert_same_events(g, o, "friction")
    steps = set(g.getPairEvents("friction")[0].tolist())
    assert steps - set(g.engine.rebuild_steps().tolist())  # pairs created between rebuilds
    u = pkg.LsmMpi(stepsPerBatch=32)
    spm.apply(u, spec)
    u.runTimeSteps(200)
    np.testing.assert_array_equal(u.getParticleData("pos"), g.getParticleData("pos"))


def test_events_of_erased_pairs(pkg, orc, spm):
    spec = falling_block(spm)
    g, o = tracked_pair(pkg, orc, spm, spec, ["friction"], 300)
    assert o.getNumParticles() < len(spec["ids"])
    assert_same_events(g, o, "friction")


def test_events_of_an_elastic_group(pkg, orc, spm):
    g, o = tracked_pair(pkg, orc, spm, settling_cloud(spm), ["elastic"], 200)
    assert_same_events(g, o, "elastic")
    assert (g.getPairEvents("elastic")[3] < 0).any()


def two_bond_groups(cls, spm, spec):
    """The block's bonds split over two RotBond groups by connection tag."""
    lsm = spm.lsm
    rest = {k: v for k, v in spec.items() if k not in ("bonds", "bond_prms", "exclusion")}
    sim = cls()
    spm.apply(sim, rest)
    b, ids, bp = spec["bonds"], spec["ids"], spec["bond_prms"]
    tags = (np.arange(len(b)) % 2).astype(np.int32)
    sim.createConnectionArrays(ids[b[:, 0]], ids[b[:, 1]], tags)
    st = bp.get("strengths")
    for t, name in ((0, "bondsA"), (1, "bondsB")):
        sim.createInteractionGroup(lsm.RotBondPrms(
            name, bp["normalK"], bp["shearK"], bp["torsionK"], bp["bendingK"], 1.0, 1.0, 1.0, 1.0,
            t, perBondStrengths=None if st is None else st[tags == t]))
        sim.createExclusion(name, "friction")
    return sim


def test_two_bonded_groups_through_erasure(pkg, orc, spm):
    """Erasure compacts every bonded group: the second group's bonds keep their
    global indices consistent (break list, bond offsets)."""
    spec = falling_block(spm)
    g = two_bond_groups(lambda: pkg.LsmMpi(stepsPerBatch=32), spm, spec)
    o = two_bond_groups(orc.oracle_lsm_class(pkg), spm, spec)
    print("case: two bonded groups, falling block, 300 steps", flush=True)
    g.runTimeSteps(300)
    o.runTimeSteps(300)
    assert o.getNumParticles() < len(spec["ids"]) and len(o.getBondBreaks()) > 50
    np.testing.assert_array_equal(g.getBondBreaks(), o.getBondBreaks())
    for name in ("bondsA", "bondsB", "friction"):
        np.testing.assert_array_equal(g.getInteractionPairs(name), o.getInteractionPairs(name))
    np.testing.assert_array_equal(g.getParticleIds(), o.getParticleIds())
    r = float(np.mean(spec["rad"]))
    assert np.max(np.abs(g.getParticleData("pos") - o.getParticleData("pos"))) <= TRAJ_TOL * r
    g.engine.check_state()


def test_checkpoint_search_that_erases(pkg, spm, tmp_path):
    """A checkpoint's neighbour search that erases particles right after a step
    whose bonds met their break criterion: the erasure keeps the stored pairs of
    the survivors (appendix pairs included) and re-indexes the pending breaks,
    which take effect at the next step as in a run without the checkpoint."""
    spec = falling_block(spm)
    ref = pkg.LsmMpi(stepsPerBatch=32)
    spm.apply(ref, spec)
    ref.runTimeSteps(300)
    brk = ref.getBondBreaks()
    # the block's lowest layer leaves the grid between the searches of steps
    # 197 and 246 (the oracle's run): a checkpoint in between erases it
    cands = sorted({int(s) - 1 for s in brk[:, 0] if 202 <= s <= 246})
    done = False
    for t in cands[:6]:
        a = pkg.LsmMpi(stepsPerBatch=32)
        spm.apply(a, spec)
        a.trackPairEvents("friction")
        a.runTimeSteps(t)
        n0 = a.getNumParticles()
        before = a.getInteractionPairs("friction")
        print("case: checkpoint at step %d" % t, flush=True)
        a.engine.save_checkpoint(str(tmp_path / ("cp%d" % t)), 0, 12)
        if a.getNumParticles() == n0:
            continue
        alive = set(a.getParticleIds().tolist())
        keep = np.array([p for p in before.tolist() if p[0] in alive and p[1] in alive])
        np.testing.assert_array_equal(a.getInteractionPairs("friction"), keep.reshape(-1, 2))
        a.engine.check_state()
        a.runTimeSteps(1)
        exp = [e for e in brk.tolist() if e[0] == t + 1 and e[1] in alive and e[2] in alive]
        got = [e for e in a.getBondBreaks().tolist() if e[0] == t + 1]
        assert got == exp
        a.runTimeSteps(300 - t - 1)
        a.engine.check_state()
        assert int(a.getPairEvents("friction")[3].sum()) == a.getNumInteractions("friction")
        done = True
        break
    assert done, "no checkpoint step among %s erased particles" % cands[:6]


def test_search_erases_before_the_break_test(pkg, orc):
    """A bond whose break criterion held in step 1 while its particle left the
    grid: step 2's search erases the particle and the bond, no break event
    (the oracle's order, the reference's); the same bond breaks when the
    particle stays."""
    tpe = importlib.import_module("test_oracle_pair_events")
    for vy, breaks in ((0.0, [[2, 0, 1]]), (-1000.0, [])):
        g = tpe.erase_break_sim(pkg.LsmMpi, pkg.lsm, vy)
        o = tpe.erase_break_sim(orc.oracle_lsm_class(pkg), pkg.lsm, vy)
        print("case: erase-vs-break vy=%g" % vy, flush=True)
        g.runTimeSteps(5)
        o.runTimeSteps(5)
        assert o.getBondBreaks().tolist() == breaks
        np.testing.assert_array_equal(g.getBondBreaks(), o.getBondBreaks())
        np.testing.assert_array_equal(g.getParticleIds(), o.getParticleIds())
        for name in ("bonded", "friction"):
            np.testing.assert_array_equal(g.getInteractionPairs(name), o.getInteractionPairs(name))
        assert np.max(np.abs(g.getParticleData("pos") - o.getParticleData("pos"))) <= 1e-12

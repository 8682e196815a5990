"""Pins tagged interaction groups (createInteractionGroupTagged ->
PIS_ED_T / PIS_NE_T, pi_storage_ed_t.hpp:15-70, pi_storage_ne_t.hpp:15-70) on
the oracle: the group holds exactly the neighbour-table pairs whose smaller tag
t1 and larger tag t2 satisfy (t1 & m1) == (T1 & m1) and (t2 & m2) == (T2 & m2),
where (T1, m1), (T2, m2) are the given (tag, mask) pairs sorted by tag."""
import numpy as np
import pytest


def cloud(pkg, orc, groups, n=400, seed=4):
    rng = np.random.default_rng(seed)
    sim = orc.oracle_lsm_class(pkg)()
    sim.initNeighbourSearch("NRotSphere", 1.2, 0.2)
    sim.setTimeStepSize(1e-4)
    sim.setSpatialDomain(pkg.BoundingBox(pkg.Vec3(0, 0, 0), pkg.Vec3(10, 10, 10)))
    tags = rng.integers(0, 6, n)
    sim.createParticleArrays(np.arange(n), rng.uniform(0.5, 9.5, (n, 3)),
                             rng.uniform(0.2, 0.5, n), tag=tags)
    for prms, t in groups:
        sim.createInteractionGroupTagged(prms, *t)
    sim.runTimeStep()
    return sim, tags


def expected(sim, tags, T1, m1, T2, m2):
    if T1 > T2:
        T1, m1, T2, m2 = T2, m2, T1, m1
    pairs = sim.engine.neighbour_pairs()
    t1 = np.minimum(tags[pairs[:, 0]], tags[pairs[:, 1]])
    t2 = np.maximum(tags[pairs[:, 0]], tags[pairs[:, 1]])
    keep = ((t1 & m1) == (T1 & m1)) & ((t2 & m2) == (T2 & m2))
    return pairs[keep]


@pytest.mark.parametrize("t", [(1, -1, 2, -1), (2, -1, 1, -1), (0, 0, 4, -1), (3, 1, 2, 6)])
def test_tagged_elastic_pairs(pkg, orc, t):
    sim, tags = cloud(pkg, orc, [(pkg.NRotElasticPrms("e", 1.0), t)])
    exp = expected(sim, tags, *t)
    got = sim.getInteractionPairs("e")
    assert len(exp) > 5
    np.testing.assert_array_equal(got, exp)


def test_tag_order_matters_with_masks(pkg, orc):
    # the pair's tags are sorted before the masked test against (0, 6), (5, -1):
    # tags (2, 5) are dropped ((2 & 6) != 0), tags (1, 5) are kept
    sim, tags = cloud(pkg, orc, [(pkg.NRotElasticPrms("e", 1.0), (5, -1, 0, 6))])
    got = sim.getInteractionPairs("e")
    t = np.sort(tags[got], axis=1)
    assert len(got) > 0
    assert np.all(t[:, 1] == 5) and np.all((t[:, 0] & 6) == 0)
    np.testing.assert_array_equal(got, expected(sim, tags, 5, -1, 0, 6))

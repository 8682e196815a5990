"""Field savers on the oracle (CLatticeMaster::runOneStep saves the fields of
step 0 before the first step and of step t after it, LatticeMaster.cpp:
2611-2632, 2672-2693):

- WallVectorFieldSaverPrms, RAW_SERIES: one line per save, each wall's vector
  as "x y z" in C++ scientific notation with 10 digits, each followed by two
  spaces (WallFieldMaster.cpp:24-44);
- InteractionScalarFieldSaverPrms, SUM: one number per save, the sum over the
  group's interactions (InteractionFieldMaster.cpp:460-487): "count" of a
  bonded group is its number of bonds; "breaking_criterion" is
  nForce/max_n + shForce/max_s + tMoment/max_t + bMoment/max_b of each bond
  (RotBondedInteraction.cpp:221-227), which for a bond pulled along its axis
  is |F| / max_n, F the force on either particle."""
import numpy as np
import pytest


def two_walls(pkg, orc, tmp_path, savers):
    sim = orc.oracle_lsm_class(pkg)()
    sim.initNeighbourSearch("RotSphere", 2.5, 0.2)
    sim.setTimeStepSize(1e-3)
    sim.setSpatialDomain(pkg.BoundingBox(pkg.Vec3(-5, -5, -5), pkg.Vec3(5, 5, 5)))
    sim.createParticleArrays([0, 1], [[0.0, -0.6, 0.0], [0.0, 0.6, 0.0]], [0.5, 0.5],
                             vel=[[0.0, -1.0, 0.0], [0.0, 1.0, 0.0]])
    sim.createWall("low", pkg.Vec3(0, -1.0, 0), pkg.Vec3(0, 1, 0))
    sim.createWall("up", pkg.Vec3(0, 1.05, 0), pkg.Vec3(0, -1, 0))
    sim.createInteractionGroup(pkg.NRotElasticWallPrms("lw", "low", 100.0))
    sim.createInteractionGroup(pkg.NRotElasticWallPrms("uw", "up", 100.0))
    sim.setWallStepMotion("up", pkg.Vec3(0.0, -0.001, 0.0))
    for s in savers:
        sim.createFieldSaver(s)
    return sim


def test_wall_raw_series(pkg, orc, tmp_path):
    ff, fp = str(tmp_path / "wf.dat"), str(tmp_path / "wp.dat")
    sim = two_walls(pkg, orc, tmp_path, [
        pkg.WallVectorFieldSaverPrms(["low", "up"], "Force", ff, "RAW_SERIES", 0, 9, 3),
        pkg.WallVectorFieldSaverPrms("up", "Position", fp, "RAW_SERIES", 2, 100, 4)])
    hist = {}
    for t in range(1, 11):
        sim.runTimeStep()
        hist[t] = (np.array(sim.getWallForce("low")), np.array(sim.getWallForce("up")),
                   np.array(sim.getWallPosition("up")))
    lines = open(ff).read().split("\n")
    assert lines[-1] == "" and len(lines) == 5  # steps 0, 3, 6, 9
    z = "0.0000000000e+00"
    assert lines[0] == ("%s %s %s  " % (z, z, z)) * 2
    for line, t in zip(lines[1:4], (3, 6, 9)):
        vals = np.array([float(x) for x in line.split()]).reshape(2, 3)
        np.testing.assert_allclose(vals[0], hist[t][0], rtol=1e-10)
        np.testing.assert_allclose(vals[1], hist[t][1], rtol=1e-10)
        assert line.endswith("  ") and "  " in line[:-2]
    assert np.any(np.abs(hist[3][0]) > 0) and np.any(np.abs(hist[9][1]) > 0)
    pos = open(fp).read().split("\n")[:-1]
    assert len(pos) == 3  # steps 2, 6, 10
    for line, t in zip(pos, (2, 6, 10)):
        np.testing.assert_allclose([float(x) for x in line.split()], hist[t][2], rtol=1e-12)
    # the up wall started at 1.05 and moved by 0.001 per step
    assert pos[0] == "0.0000000000e+00 1.0480000000e+00 0.0000000000e+00  "


def test_bond_count_and_criterion_sum(pkg, orc, tmp_path):
    fc, fk = str(tmp_path / "nb.dat"), str(tmp_path / "crit.dat")
    sim = orc.oracle_lsm_class(pkg)()
    sim.initNeighbourSearch("RotSphere", 2.5, 0.2)
    sim.setTimeStepSize(1e-4)
    sim.setSpatialDomain(pkg.BoundingBox(pkg.Vec3(-5, -5, -5), pkg.Vec3(5, 5, 5)))
    # three bonded pairs pulled apart along x; the last one breaks first
    xs = [[-3.0, 0, 0], [-2.0, 0, 0], [0.0, 0, 0], [1.0, 0, 0], [3.0, 0, 0], [4.0, 0, 0]]
    v = [[0, 0, 0], [0.5, 0, 0], [0, 0, 0], [0.5, 0, 0], [0, 0, 0], [2.0, 0, 0]]
    sim.createParticleArrays(range(6), xs, [0.5] * 6, vel=v)
    sim.createConnectionArrays([0, 2, 4], [1, 3, 5], [0, 0, 0])
    sim.createInteractionGroup(pkg.RotBondPrms("b", 100.0, 50.0, 50.0, 50.0, 2.0, 1e30, 1e30, 1e30,
                                               0, scaling=False))
    sim.createFieldSaver(pkg.InteractionScalarFieldSaverPrms("b", "count", fc, "SUM", 0, 1000, 10))
    sim.createFieldSaver(pkg.InteractionScalarFieldSaverPrms("b", "breaking_criterion", fk, "SUM",
                                                             1, 1000, 10))
    forces, counts = {}, {}
    for t in range(1, 121):
        sim.runTimeStep()
        f = sim.getParticleData("force")
        forces[t] = [np.linalg.norm(f[1]), np.linalg.norm(f[3]), np.linalg.norm(f[5])]
        counts[t] = sim.getNumInteractions("b")
    nb = [float(x) for x in open(fc).read().split()]
    assert nb[0] == 3.0 and len(nb) == 13  # steps 0, 10, ..., 120
    assert nb == [3.0] + [float(counts[t]) for t in range(10, 121, 10)]
    assert nb[-1] < 3.0  # a bond broke and left the count
    crit = [float(x) for x in open(fk).read().split()]
    assert len(crit) == 12  # steps 1, 11, ..., 111
    for c, t in zip(crit, range(1, 121, 10)):
        # a removed bond no longer pulls its particle, so the sum runs over all three
        np.testing.assert_allclose(c, sum(forces[t]) / 2.0, rtol=1e-9, atol=1e-15)
    assert max(crit) > 0.5
    assert open(fc).read().split("\n")[0] == "3.0000000000e+00"


def test_saver_parameters(pkg):
    s = pkg.InteractionScalarFieldSaverPrms("b", "count", "x", "SUM", 3, 10, 3)
    assert [t for t in range(0, 15) if s.needSave(t)] == [3, 6, 9]
    assert s.nextSave(0) == 3 and s.nextSave(3) == 6 and s.nextSave(9) is None
    with pytest.raises(ValueError):
        pkg.WallVectorFieldSaverPrms("w", "Velocity", "x", "RAW_SERIES", 0, 1, 1)
    with pytest.raises(ValueError):
        pkg.InteractionScalarFieldSaverPrms("b", "count", "x", "RAW_SERIES", 0, 1, 1)

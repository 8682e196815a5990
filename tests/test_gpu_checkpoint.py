"""Restart checkpoints in the reference's text format (CheckPointInfo.cpp:73-96,
SubLattice.hpp:3599-3790): the files hold the engine's state record by record,
save -> load -> save reproduces the files, a run restarted from a checkpoint
follows the run that saved it, and the reader takes the RotFriction section as
the reference writes it (count and first record run together)."""
import os

import numpy as np
import pytest

from test_gpu_parity import TRAJ_TOL, lattice_spec, rel_field

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def spm(pkg):
    import importlib
    return importlib.import_module("esys_particle_amd.specimens")


def setup_groups(pkg, sim, spec):
    """The script of a restart: model, groups and walls, no particles or connections."""
    sim.initNeighbourSearch(spec["ptype"], spec["grid"], spec["alpha"])
    sim.setTimeStepSize(spec["dt"])
    lo, hi = spec["domain"]
    sim.setSpatialDomain(pkg.BoundingBox(pkg.Vec3(*lo), pkg.Vec3(*hi)))
    bp = spec["bond_prms"]
    sim.createInteractionGroup(pkg.RotBondPrms(bp["name"], bp["normalK"], bp["shearK"],
                                               bp["torsionK"], bp["bendingK"], 1.0, 1.0, 1.0,
                                               1.0, bp["tag"]))
    name, kn, mud, mus, kss = spec["friction"]
    sim.createInteractionGroup(pkg.RotFrictionPrms(name, kn, mud, mus, kss))
    sim.createExclusion(*spec["exclusion"])
    for wname, o, nrm, k, step in spec["walls"]:
        sim.createWall(wname, pkg.Vec3(*o), pkg.Vec3(*nrm))
        sim.createInteractionGroup(pkg.NRotElasticWallPrms(wname + "_wall", wname, k))
        if any(step):
            sim.setWallStepMotion(wname, pkg.Vec3(*step))
    return sim


def saved(sim, prefix, precision=12):
    sim.engine.save_checkpoint(prefix, 0, precision)
    t = sim.getTimeStep()
    return "%s_t=%d_0.txt" % (prefix, t), "%s_t=%d_1.txt" % (prefix, t)


def sections(path):
    tok = open(path).read().split()
    k = 6
    n = int(tok[k]); k += 1
    parts = []
    for _ in range(n):
        parts.append(tok[k:k + 32]); k += 32
    nbg = int(tok[k]); k += 1
    bonds = []
    for _ in range(nbg):
        assert tok[k] == "RotBonded"
        nb = int(tok[k + 1]); k += 2
        bonds.append([tok[k + 20 * b: k + 20 * (b + 1)] for b in range(nb)]); k += 20 * nb
    ndg = int(tok[k]); k += 1
    fric = []
    for _ in range(ndg):
        assert tok[k] == "RotFriction"
        nf = int(tok[k + 1]); k += 2
        fric.append([tok[k + 15 * c: k + 15 * (c + 1)] for c in range(nf)]); k += 15 * nf
    assert tok[k] == "Walls"
    nw = int(tok[k + 1]); k += 2
    walls = {tok[k + 10 * w]: [float(v) for v in tok[k + 10 * w + 1: k + 10 * (w + 1)]]
             for w in range(nw)}
    k += 10 * nw
    assert tok[k:] == ["Spheres", "0", "TriMesh", "0", "Mesh2D", "0"]
    return parts, bonds, fric, walls


def test_checkpoint_files_hold_the_state(pkg, spm, tmp_path):
    spec = lattice_spec(spm)
    g = pkg.LsmMpi()
    spm.apply(g, spec)
    g.runTimeSteps(40)
    info, data = saved(g, str(tmp_path / "cp"), precision=17)
    itok = open(info).read().split()
    assert itok[:2] == ["V", "3"] and int(itok[4]) == 40
    assert itok[5:7] == ["ESyS-Particle_2.3.5_bzr_rev.", "1218"]
    assert itok[7] == "BoundingBox" and itok[14] == "PeriodicBoundaries" and itok[18] == "Dimension"
    assert os.path.basename(itok[20]) == os.path.basename(data)
    parts, bonds, fric, walls = sections(data)
    ids = g.getParticleIds()
    assert [int(p[4]) for p in parts] == list(ids)
    P = np.array([[float(v) for v in p[0:3]] for p in parts])
    np.testing.assert_array_equal(P, g.getParticleData("pos"))
    np.testing.assert_array_equal(np.array([[float(v) for v in p[13:16]] for p in parts]),
                                  g.getParticleData("vel"))
    np.testing.assert_array_equal(np.array([[float(v) for v in p[24:28]] for p in parts]),
                                  g.getParticleData("quat"))
    # the forced search before the checkpoint made every old position current
    np.testing.assert_array_equal(np.array([[float(v) for v in p[10:13]] for p in parts]), P)
    pairs = sorted((int(b[0]), int(b[1])) for b in bonds[0])
    assert pairs == sorted(map(tuple, g.getInteractionPairs("bonded").tolist()))
    fp = [(int(c[8]), int(c[9])) for c in fric[0]]
    assert fp == list(map(tuple, g.getInteractionPairs("friction").tolist()))
    ff, fl = g.engine.friction_state("friction")
    np.testing.assert_array_equal(np.array([[float(v) for v in c[12:15]] for c in fric[0]]), ff)
    np.testing.assert_array_equal([int(c[10]) | (int(c[11]) << 1) for c in fric[0]], fl)
    for w in ("bottom", "top"):
        np.testing.assert_array_equal(walls[w][0:3], list(g.getWallPosition(w)))


def test_save_load_save_reproduces_files(pkg, spm, tmp_path):
    spec = lattice_spec(spm)
    a = pkg.LsmMpi()
    spm.apply(a, spec)
    info, data = saved(a, str(tmp_path / "a"))
    b = setup_groups(pkg, pkg.LsmMpi(), spec)
    b.loadCheckPoint(info)
    info2, data2 = saved(b, str(tmp_path / "b"))
    assert open(data).read() == open(data2).read()


@pytest.mark.parametrize("layout", ["separated", "reference"])
def test_restart_follows_the_saving_run(pkg, spm, tmp_path, layout):
    spec = lattice_spec(spm)
    a = pkg.LsmMpi()
    spm.apply(a, spec)
    # a step after which no bond is pending (no break event at step 62): the
    # checkpoint does not hold the cached bond forces, so a bond whose criterion
    # held at the saved step would break one step later after a restart, in the
    # reference as here
    a.runTimeSteps(61)
    info, data = saved(a, str(tmp_path / "cp"), precision=17)
    if layout == "reference":
        # the reference's PIS_ED writer puts the first record right after the count
        txt = open(data).read()
        head, rest = txt.split("RotFriction\n", 1)
        count, tail = rest.split("\n", 1)
        assert int(count) > 0
        open(data, "w").write(head + "RotFriction\n" + count + tail)
    a.runTimeSteps(60)
    b = setup_groups(pkg, pkg.LsmMpi(), spec)
    b.loadCheckPoint(info)
    b.runTimeSteps(60)
    assert b.getTimeStep() == 121
    assert 62 not in a.getBondBreaks()[:, 0]
    r = float(np.mean(spec["rad"]))
    assert np.max(np.abs(b.getParticleData("pos") - a.getParticleData("pos"))) <= TRAJ_TOL * r
    assert rel_field(b.getParticleData("vel"), a.getParticleData("vel")) <= TRAJ_TOL
    ea = a.getBondBreaks()
    np.testing.assert_array_equal(b.getBondBreaks(), ea[ea[:, 0] > 61])
    np.testing.assert_array_equal(b.getInteractionPairs("bonded"), a.getInteractionPairs("bonded"))
    np.testing.assert_array_equal(b.getInteractionPairs("friction"),
                                  a.getInteractionPairs("friction"))

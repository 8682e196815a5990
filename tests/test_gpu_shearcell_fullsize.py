"""The frictional shear cell at its full size, 2,000,000 rotational spheres,
periodic in x (random sequential packing), rough bonded walls sliding at
+-0.1 and the servo-held confining stress: size-independent properties of the
HIP engine's run, where the oracle would take too long.

- the neighbour set of the first build equals a brute-force kd-tree search
  with the nearest x image and the exact predicate
  |x_i - x_j|^2 <= (r_i + r_j + alpha)^2 (ntable.hpp:471-521), and the
  friction group holds exactly those pairs (no exclusions);
- after 100 steps: every force is finite; the particles' total force equals
  the walls' forces (a wall's force is the sum of the forces its interactions
  apply to particles, Wall.h:68 with EWallInteraction.hpp:75-76 and
  BWallInteraction.hpp:46-47; pair forces cancel, so sum_i F_i = sum_walls W to
  rounding); the servo holds the top wall's normal force at the stress times
  the area; positions stay within one Verlet skin of the circular x range."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_shear_cell_2m_properties(pkg):
    from scipy.spatial import cKDTree
    spm = importlib.import_module("esys_particle_amd.specimens")
    spec = spm.shear_cell(n=2000000, seed=11)
    n = len(spec["ids"])
    L = spec["box"][0]
    alpha = spec["alpha"]
    print("case: shear cell n=%d, box %s" % (n, spec["box"]), flush=True)
    sim = pkg.LsmMpi(stepsPerBatch=32)
    spm.apply(sim, spec)
    sim.runTimeSteps(1)
    # the first build ran on the initial positions
    pos0, rad = spec["pos"], spec["rad"]
    got = sim.engine.neighbour_pairs()
    tree = cKDTree(np.c_[np.mod(pos0[:, 0], L), pos0[:, 1], pos0[:, 2]], boxsize=[L, 1e9, 1e9])
    cand = tree.query_pairs(2 * rad.max() + alpha, output_type="ndarray")
    d = pos0[cand[:, 0]] - pos0[cand[:, 1]]
    d[:, 0] -= L * np.round(d[:, 0] / L)
    lim = rad[cand[:, 0]] + rad[cand[:, 1]] + alpha
    keep = (d * d).sum(1) <= lim * lim
    ids = spec["ids"]
    a, b = ids[cand[keep, 0]], ids[cand[keep, 1]]
    want = np.unique(np.minimum(a, b).astype(np.int64) << 32 | np.maximum(a, b).astype(np.int64))
    have = np.unique(got[:, 0].astype(np.int64) << 32 | got[:, 1].astype(np.int64))
    print("neighbour pairs: engine %d, brute force %d" % (len(have), len(want)), flush=True)
    np.testing.assert_array_equal(have, want)
    fr = sim.getInteractionPairs("friction")
    np.testing.assert_array_equal(
        np.unique(fr[:, 0].astype(np.int64) << 32 | fr[:, 1].astype(np.int64)), want)
    sim.runTimeSteps(99)
    f = sim.getParticleData("force")
    assert np.all(np.isfinite(f)) and np.all(np.isfinite(sim.getParticleData("pos")))
    walls = np.array([list(sim.getWallForce(w)) for w in ("top", "bottom")])
    scale = np.abs(f).sum(0).max()
    balance = f.sum(0) - walls.sum(0)
    print("force balance %s (scale %.3g)" % (balance, scale), flush=True)
    assert np.all(np.abs(balance) <= 1e-8 * scale)
    area = spec["box"][0] * spec["box"][2]
    assert abs(walls[0][1] + 1.0 * area) < 1e-2 * area
    x = sim.getParticleData("pos")[:, 0]
    assert np.all((x > -alpha) & (x < L + alpha))

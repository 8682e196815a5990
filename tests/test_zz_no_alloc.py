"""esp_apply_force_to_wall_noalloc, the allocation-free servo (include/esp.h),
against esp_apply_force_to_wall: called every time step on a periodic shear
cell with a servo-loaded bonded wall and, interleaved with it, on a smaller
compression specimen with a servo-loaded elastic wall (so each call follows a
call with more candidates on the same GPU).  Wall positions and particle
states equal bit for bit."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sim(pkg, spm, spec):
    sim = pkg.LsmMpi(stepsPerBatch=32)
    spm.apply(sim, spec)
    return sim


def test_noalloc_servo_matches_allocating_servo(pkg):
    spm = importlib.import_module("esys_particle_amd.specimens")
    rng = np.random.default_rng(3)
    big = spm.shear_cell(n=1500, seed=3, phi=0.35, dt=5e-4, wall_speed=10.0)
    big["vel"] = rng.normal(0.0, 2.0, (1500, 3))
    g_big, f_big = big["servo"]
    big["servo"] = None
    small = spm.bonded_compression(n=600, seed=4, coh_mean=5.0, wall_speed=2.0)
    area = small["box"][0] * small["box"][2] if "box" in small else 1.0
    g_small, f_small = "top_wall", (0.0, -50.0 * area, 0.0)
    a, b = _sim(pkg, spm, big), _sim(pkg, spm, big)        # big: allocating / allocation-free
    c, d = _sim(pkg, spm, small), _sim(pkg, spm, small)    # small: allocation-free / allocating
    print("case: servo big n=1500 (bonded wall) then small n=600 (elastic wall), 80 steps",
          flush=True)
    for t in range(80):
        if t >= 1:
            a.engine.apply_force_to_wall(g_big, f_big)
            b.engine.apply_force_to_wall_noalloc(g_big, f_big)
            c.engine.apply_force_to_wall_noalloc(g_small, f_small)
            d.engine.apply_force_to_wall(g_small, f_small)
        for s in (a, b, c, d):
            s.runTimeSteps(1)
    for x, y, walls in ((a, b, ("top", "bottom")), (d, c, ("top", "bottom"))):
        for w in walls:
            assert list(x.getWallPosition(w)) == list(y.getWallPosition(w))
        for f in ("pos", "vel", "angvel"):
            np.testing.assert_array_equal(x.getParticleData(f), y.getParticleData(f))
    # the servos moved their walls
    assert list(a.getWallPosition("top")) != [0.0, big["box"][1], 0.0]

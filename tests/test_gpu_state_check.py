"""The device bookkeeping stays consistent through long queued runs
(esp_check_state after every batch): per-particle appendix slot counts, the
flagged-particle list and event count, every stored pair and bond with its
force-record slots inside its particles' summation runs, appendix counts.
Runs that exercise all of it: a dense bonded specimen through its bond-break
cascade (flagged pairs appended on the device, halts for the appendix, host
phases, neighbour rebuilds), a bonded block falling out of the grid (erasure),
a periodic shear cell under the servo."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def spm(pkg):
    return importlib.import_module("esys_particle_amd.specimens")


def test_dense_bonded_run_bookkeeping(pkg, spm):
    spec = spm.dense_bonded_compression(n=20000, seed=3)
    sim = pkg.LsmMpi(stepsPerBatch=128)
    spm.apply(sim, spec)
    print("case: dense bonded 20000, 3200 steps in batches of 128", flush=True)
    for _ in range(25):
        sim.runTimeSteps(128)
        sim.engine.check_state()
    st = sim.engine.stats()
    print(st, flush=True)
    assert st["num_broken"] > 1000 and st["num_host_phases"] >= 2


def test_erasure_and_periodic_bookkeeping(pkg, spm):
    spec = spm.lattice_block(nx=6, ny=10, nz=6, dt=1e-4, vel_sd=0.05)
    spec["vel"] = spec["vel"] + np.array([0.0, -20.0, 0.0])
    spec["walls"] = []
    spec["gravity"] = (0.0, -9.81, 0.0)
    spec["domain"] = ((-1.0, 2.4, -1.0), spec["domain"][1])
    sim = pkg.LsmMpi(stepsPerBatch=32)
    spm.apply(sim, spec)
    print("case: bonded block erased outside the grid, 300 steps", flush=True)
    for _ in range(12):
        sim.runTimeSteps(25)
        sim.engine.check_state()
    assert sim.getNumParticles() < 340
    spec = spm.shear_cell(n=1500, seed=3, phi=0.35, dt=5e-4, wall_speed=10.0)
    spec["vel"] = np.random.default_rng(3).normal(0.0, 2.0, (1500, 3))
    sim = pkg.LsmMpi(stepsPerBatch=32)
    spm.apply(sim, spec)
    print("case: periodic shear cell 1500 with the servo, 300 steps", flush=True)
    for _ in range(6):
        sim.runTimeSteps(50)
        sim.engine.check_state()
    assert len(sim.engine.rebuild_steps()) >= 5

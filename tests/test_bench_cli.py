"""bench.py's contract on the CPU: one GPU only (--gpus N>1 exits 2 with a
message), the per-unit byte counts the roofline uses, and the CPU baseline
leg (the oracle stepping a bounded sample of the specimen)."""
import importlib
import os
import subprocess
import sys

from conftest import ROOT, _load


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_algorithmic_bytes_per_unit():
    bench = _load("esp_bench", os.path.join(ROOT, "bench.py"))
    # FP64 state: rot particle reads 19 doubles, writes 13; non-rot 11 and 6
    assert bench.B_PARTICLE_ROT == (19 + 13) * 8
    assert bench.B_PARTICLE_NROT == (11 + 6) * 8
    # bond: two int32 ids + 4 stiffnesses + 4 strengths + D0 (3 doubles)
    assert bench.B_BOND == 2 * 4 + (4 + 4 + 3) * 8
    # friction: two ids + shear spring (3 doubles) + flags (int32), read and written
    assert bench.B_FRICTION == 8 + 2 * (3 * 8 + 4)


def test_cpu_baseline_leg(pkg):
    bench = _load("esp_bench", os.path.join(ROOT, "bench.py"))
    spm = importlib.import_module("esys_particle_amd.specimens")
    spec = spm.lattice_block(nx=4, ny=6, nz=4, dt=1e-4, vel_sd=0.05, wall_speed=5.0)
    out, steps = bench.cpu_baseline(pkg, spm, spec, budget_s=0.3)
    assert out["unit"] == "particle-steps/s" and out["kind"] == "port" and out["cores"] == 1
    assert out["value"] > 0 and steps >= 10 and steps % 10 == 0
    assert out["sample"].startswith("steps 2-%d of the same specimen, " % (steps + 1))


def test_cpu_baseline_all_cores_leg(pkg, tmp_path):
    """Independent single-thread oracle processes on copies of the specimen;
    the aggregate rate, the process count and the CPU model are reported."""
    bench = _load("esp_bench", os.path.join(ROOT, "bench.py"))
    spm = importlib.import_module("esys_particle_amd.specimens")
    spec = spm.lattice_block(nx=4, ny=6, nz=4, dt=1e-4, vel_sd=0.05, wall_speed=5.0)
    path = str(tmp_path / "spec.pkl")
    bench.save_specimen(path, spec)
    out = bench.cpu_baseline_all_cores(path, 0.3, 2)
    assert out["cores"] == 2 and out["kind"] == "port" and out["value"] > 0
    assert out["nproc"] >= 1 and isinstance(out["cpu_model"], str)


def test_idle_friction_pair_bytes():
    """A stored friction pair whose particles are apart needs its two ids and
    the flag word that says its history is zero."""
    bench = _load("esp_bench", os.path.join(ROOT, "bench.py"))
    assert bench.B_FRICTION_IDLE == 8 + 4

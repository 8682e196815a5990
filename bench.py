"""Benchmark: bonded rotational uniaxial compression, 100,000 spheres, FP64.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Workload (one step = one runTimeStep of the whole specimen on one GPU):
  100,000 RotSphere, radii log-uniform in [0.2, 1.0], dense seeded packing in a
  1:2:1 prism; bonds (RotBondPrms kr=1e5, ks=kt=kb=3e4) for pairs with
  |x_i - x_j| <= 1.01 (r_i + r_j), per-bond tensile/shear cohesion ~ N(50, 5);
  RotFriction mu=0.6 on unbonded contacts (excluded while bonded); top and bottom
  elastic walls approaching at 0.05 per unit time; dt = 1e-5.

Prints one JSON line: throughput in particle-steps/s over the timed steps, the
roofline of the dominant kernel (algorithmic bytes / HIP-event-measured launch
time against 8 TB/s), and the oracle (CPU restatement) timed on a bounded sample
of the same specimen, on one core and as one process per core of the CPU share.
"""
import argparse
import importlib
import importlib.util
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0

# algorithmic bytes (compulsory state traffic) per unit, FP64 structure of arrays
B_PARTICLE_ROT = 256  # read pos3 vel3 w3 q4 r 1/m 1/I old3 (19 doubles), write pos vel w q (13)
B_PARTICLE_NROT = 136  # read pos3 vel3 r 1/m old3 (11 doubles), write pos vel (6)
B_BOND = 96           # i, j (2 x int32) + kr ks kt kb + 4 strengths + D0 (3 doubles)
B_FRICTION = 64       # touching contact: i, j + shear spring (3) + flags, read and write
B_FRICTION_IDLE = 12  # stored pair apart: i, j + the flag word saying its history is zero
B_ELASTIC = 8         # i, j

# FP64 work per unit, from the SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 counters of the
# default workload (64 lanes x (ADD + MUL + 2 FMA + TRANS) per dispatch over its units)
FLOP_BOND = 1386            # bond law + break test (k_forces)
FLOP_PARTICLE_GATHER = 189  # rotational particle: slot sums, walls, integration
FP64_PEAK_GFLOPS = 78600.0
# HBM bytes per launch measured with FETCH_SIZE / WRITE_SIZE (copy-probe scaled) for
# the default workload, in profiles/step_kernels_bonded100k.md
MEASURED_TRAFFIC = {("compression", 100000): {"forces": 60.6e6, "gather_integrate": 57.8e6}}


def _load(name, path, pkgdir=None):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, path, submodule_search_locations=[pkgdir] if pkgdir else None)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def load_pkg():
    d = os.path.join(ROOT, "esys-particle_amd")
    return _load("esys_particle_amd", os.path.join(d, "__init__.py"), d)


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def copy_probe(mib):
    """A plain device-to-device copy of known size, for scaling the FETCH_SIZE /
    WRITE_SIZE counters (each copy reads and writes mib MiB)."""
    import torch
    a = torch.ones(mib << 20, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(10):
        b.copy_(a)
    torch.cuda.synchronize()
    log("copy probe: 10 x %d MiB" % mib)


def save_specimen(path, spec):
    """Cache of a specimen this script built (its own file, read back below)."""
    import pickle
    with open(path, "wb") as f:
        pickle.dump(spec, f)


def load_specimen(path):
    import pickle
    with open(path, "rb") as f:
        return pickle.load(f)


def cpu_baseline(pkg, spm, spec, budget_s, max_steps=0):
    """The oracle (C restatement of the reference path, one thread) stepping the
    same specimen from its initial state until ~budget_s seconds have passed or
    (max_steps > 0) it reached the last step the timed GPU pass ran, so the
    sample never runs past the trajectory the GPU figure covers."""
    orc = _load("esp_oracle_driver", os.path.join(ROOT, "oracle", "oracle.py"))
    sim = orc.oracle_lsm_class(pkg)()
    spm.apply(sim, spec)
    sim.runTimeSteps(1)  # first step: initial neighbour search, excluded
    chunk = 10 if sim.getNumParticles() <= 200000 else 1
    steps, psteps, t0 = 0, 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s and (max_steps <= 0 or steps + 1 < max_steps):
        if max_steps > 0:
            chunk = min(chunk, max_steps - steps - 1)
        sim.runTimeSteps(chunk)
        steps += chunk
        psteps += chunk * sim.getNumParticles()
    el = time.perf_counter() - t0
    return {"value": psteps / el, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": "steps 2-%d of the same specimen, %d particles at the end (%.1f s)" % (
                steps + 1, sim.getNumParticles(), el)}, steps


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_all_cores(spec_path, budget_s, workers, max_steps=0):
    """`workers` single-thread oracle processes, each stepping its own copy of
    the same specimen for ~budget_s seconds: aggregate particle-steps/s, the
    throughput of the host cores on independent copies (an upper bound on a
    domain-decomposed run of the reference over the same cores).  Child
    processes, so nothing here shares the GPU process's state."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-worker", spec_path,
           "--cpu-budget", str(budget_s), "--cpu-max-steps", str(max_steps)]
    procs = [subprocess.Popen(cmd, stdout=subprocess.PIPE, text=True) for _ in range(workers)]
    res = []
    for p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError("cpu worker failed (%d)" % p.returncode)
        res.append(json.loads(out.strip().splitlines()[-1]))
    return {"value": sum(r["value"] for r in res), "unit": "particle-steps/s",
            "cores": workers, "kind": "port", "nproc": os.cpu_count(),
            "cpu_model": cpu_model(),
            "sample": "%d concurrent single-thread oracle processes, each steps 2-~%d of its "
                      "own copy of the same specimen (%.0f s each)" % (
                          workers, min(r["steps"] for r in res) + 1, budget_s)}


def condition(sim, eng, steps, settling):
    """Specimen preparation before the warm-up: the compression specimen starts
    with a bond-break cascade (bonds whose initial gap exceeds their failure
    strain break in the first steps, ~23,000 of them at 100,000 particles) whose
    flagged-pair list holds ~1.2 million friction pairs until the next neighbour
    rebuild.  steps < 0: run until that rebuild has happened (the steady state
    that makes up the rest of a 20,000-step run), at most 5,000 steps; 0: none.
    Untimed, like the warm-up; returns the number of steps run."""
    if steps == 0 or settling:
        return 0
    t0 = time.perf_counter()
    if steps > 0:
        sim.runTimeSteps(steps)
        done = steps
    else:
        sim.runTimeSteps(1)
        done = 1
        first = eng.stats()["num_rebuilds"]
        while done < 5000 and eng.stats()["num_rebuilds"] == first:
            sim.runTimeSteps(100)
            done += 100
    log("conditioning %d steps: %.2f s, stats %s" % (done, time.perf_counter() - t0, eng.stats()))
    return done


def calibration_bench(args):
    """The batched calibration sweep: every specimen's loss (mean squared
    difference of its sampled axial stress to a target curve) and gradient with
    respect to (k_n, k_s, cohesion, mu) from one batched forward run and one
    adjoint sweep; value = specimens / wall time of both (gradient evaluations
    per second).  Baseline: central finite differences on the plain engine,
    2 runs per parameter, each specimen alone, timed on a sample of specimens
    across the size range (the faster of two runs each) and summed over the
    batch by interpolation in N."""
    pkg = load_pkg()
    spm = importlib.import_module("esys_particle_amd.specimens")
    cal = importlib.import_module("esys_particle_amd.calibration")
    t0 = time.perf_counter()
    specs, theta = spm.calibration_batch(nspec=args.specimens, workers=1, progress=log)
    ns = np.array([len(sp["rad"]) for sp in specs])
    log("specimens: %d, %d particles, %d bonds (%.1f s)" % (
        len(specs), ns.sum(), sum(len(sp["bonds"]) for sp in specs), time.perf_counter() - t0))
    every = 20
    tune = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        tune = cal.autotune(specs, theta, trial_steps=60, sample_every=every, log=log)
        log("autotune: %.1f s -> block %d alpha %.3g grid %.4g" % (
            time.perf_counter() - t0, tune["block"], tune["alpha"], tune["grid"]))
    kw = {} if tune is None else {"block": tune["block"], "alpha": tune["alpha"],
                                  "grid": tune["grid"], "gap": tune["gap"]}
    steps = args.steps
    # warm-up: code loading, allocations
    mw = cal.DiffModel.from_specimens(specs, max(args.warmup, 20), every, **kw)
    mw.forward(theta, np.zeros((len(specs), mw.nsamples)))
    mw.backward()
    mw.close()
    m = cal.DiffModel.from_specimens(specs, steps, every, checkpoint_every=100, **kw)
    target = np.zeros((len(specs), m.nsamples))
    t0 = time.perf_counter()
    _, loss = m.forward(theta, target)
    tf = time.perf_counter() - t0
    t1 = time.perf_counter()
    grad = m.backward()
    tb = time.perf_counter() - t1
    fwd_ms, bwd_ms = m.times_ms()
    value = len(specs) / (tf + tb)
    log("forward %.2f s (device %.0f ms), adjoint %.2f s (device %.0f ms), %d breaks, "
        "rebuilds %s" % (tf, fwd_ms, tb, bwd_ms, len(m.bond_breaks()), m.rebuild_steps().tolist()))
    m.close()
    # plain engine, one specimen at a time, forward runs of the same length
    order = np.argsort(ns)
    pick = order[np.linspace(0, len(order) - 1, min(6, len(order))).astype(int)]
    # each sample timed twice, the faster run kept (the baseline's runs are
    # launch-bound and vary between runs; the faster one is the conservative
    # baseline for the ratio)
    pts = []
    for s in pick:
        best = None
        for _ in range(2):
            mm = cal.DiffModel.from_specimens([specs[s]], steps, every)
            sim = pkg.LsmMpi(stepsPerBatch=128)
            cal.plain_setup(sim, mm, theta[s])
            sim.runTimeSteps(1)
            t0 = time.perf_counter()
            sim.runTimeSteps(steps - 1)
            el = time.perf_counter() - t0
            mm.close()
            best = el if best is None else min(best, el)
        pts.append((int(ns[s]), best * steps / max(steps - 1, 1)))
        log("plain engine: %d particles, %d steps: %.3f s (faster of 2)" % (pts[-1][0], steps,
                                                                          pts[-1][1]))
    px, py = np.array([p[0] for p in pts], float), np.array([p[1] for p in pts], float)
    t_plain = float(np.interp(ns, px, py).sum())
    nparam = 4
    fd_value = len(specs) / (2 * nparam * t_plain)
    out = {
        "metric": "gradient evaluations/s",
        "value": value,
        "unit": "gradients/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": (tf + tb) / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "batched calibration sweep: %d bonded compression specimens, "
                        "N log-uniform [2000, 30000], per-specimen k_n in [5e4, 2e5], "
                        "k_s = 0.3 k_n, cohesion in [20, 80], mu in [0.3, 0.8]; forward "
                        "%d steps + adjoint, stress sampled every %d steps, checkpoints "
                        "every 100 steps" % (len(specs), steps, every),
            "specimens": len(specs),
            "particles": int(ns.sum()),
            "autotune": None if tune is None else {k: tune[k] for k in ("block", "alpha", "grid")},
        },
        "forward_s": tf,
        "adjoint_s": tb,
        "fd_baseline": {"value": fd_value, "unit": "gradients/s",
                        "runs_per_gradient": 2 * nparam,
                        "plain_forward_s": pts,
                        "ratio": value / fd_value},
        "loss_mean": float(np.mean(loss)),
        "grad_finite": bool(np.all(np.isfinite(grad))),
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: 20000; calibration: 1000, the sweep's run length)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed steps before them (default: 500; calibration: 20)")
    ap.add_argument("--particles", type=int, default=100000)
    ap.add_argument("--batch", type=int, default=128, help="steps queued per host sync")
    ap.add_argument("--condition", type=int, default=-1,
                    help="compression: steps run before the warm-up to take the specimen "
                         "through its initial bond-break cascade (-1: until the first "
                         "neighbour rebuild after it, at most 5000; 0: none)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--cpu-workers", type=int, default=0,
                    help="all-cores CPU leg: this many oracle processes (0: the CPU share, "
                         "at most 16)")
    ap.add_argument("--cpu-worker", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-max-steps", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--specimen", default=None,
                    help="npz file: load the specimen from it if present, else build and save it")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device 10 times (counter calibration)")
    ap.add_argument("--workload", default="compression",
                    choices=["compression", "settling", "shearcell", "calibration"],
                    help="compression: the bonded 100k-particle compression (default); "
                         "settling: 10,000 NRotSphere settling under gravity on an elastic "
                         "floor (NRotElastic k=1, dt=1e-4, the reference's CPU-sized case); "
                         "shearcell: 2,000,000 RotSphere frictional shear cell, periodic in "
                         "x, rough walls sliding at +-0.1, servo-held confining stress 1.0; "
                         "calibration: the batched forward + adjoint sweep over --specimens "
                         "bonded specimens of 2,000-30,000 particles, --steps steps each")
    ap.add_argument("--specimens", type=int, default=256,
                    help="calibration: specimens in the batch")
    ap.add_argument("--no-autotune", action="store_true",
                    help="calibration: default launch / search settings")
    args = ap.parse_args()
    calib = args.workload == "calibration"
    if args.steps is None:
        args.steps = 1000 if calib else 20000
    if args.warmup is None:
        args.warmup = 20 if calib else 500
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine holds one subdomain on one device "
              "(runs over several GPUs are not supported)", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.workload == "calibration":
        calibration_bench(args)
        return
    if args.cpu_worker:
        pkg = load_pkg()
        spm = importlib.import_module("esys_particle_amd.specimens")
        r, steps = cpu_baseline(pkg, spm, load_specimen(args.cpu_worker), args.cpu_budget,
                                args.cpu_max_steps)
        r["steps"] = steps
        print(json.dumps(r), flush=True)
        return

    pkg = load_pkg()
    spm = importlib.import_module("esys_particle_amd.specimens")
    t0 = time.perf_counter()
    settling = args.workload == "settling"
    shear = args.workload == "shearcell"
    spec = load_specimen(args.specimen) if args.specimen and os.path.exists(args.specimen) else None
    if spec is None:
        if shear:
            # 2,000,000 spheres, radii U[0.25, 0.75]: the solid volume (1.3e6) exceeds the
            # stated 200x40x100 box, so the box keeps its aspect at solid fraction 0.58
            # (grow-and-relax packing, periodic in x)
            spec = spm.shear_cell(n=args.particles if args.particles != 100000 else 2000000,
                                  seed=11, phi=0.58, packing="dense", sweeps=600,
                                  progress=log)
        elif settling:
            # the soft floor (k=1) lets the lowest particles sink through it and
            # out of the neighbour table, which erases them, as the reference does
            spec = spm.gravity_settling(n=10000, seed=1, box=(20.0, 40.0, 20.0), dt=1e-4)
        else:
            spec = spm.dense_bonded_compression(n=args.particles, seed=3)
        if args.specimen:
            save_specimen(args.specimen, spec)
    log("specimen: %d particles, %d bonds (%.1f s)" % (
        len(spec["ids"]), len(spec.get("bonds", [])), time.perf_counter() - t0))

    sim = pkg.LsmMpi(device=0, stepsPerBatch=args.batch)
    spm.apply(sim, spec)
    eng = sim.engine
    cond = condition(sim, eng, args.condition, settling or shear)
    t0 = time.perf_counter()
    sim.runTimeSteps(args.warmup)
    log("warmup %d steps: %.2f s, stats %s" % (args.warmup, time.perf_counter() - t0, eng.stats()))

    s0 = eng.stats()
    t0 = time.perf_counter()
    sim.runTimeSteps(args.steps)  # returns after the last step completed on the device
    el = time.perf_counter() - t0
    s1 = eng.stats()
    n = s1["num_particles"]
    value = n * args.steps / el
    log("timed %d steps: %.3f s, stats %s" % (args.steps, el, s1))

    # per-kernel device time, HIP events on the engine's stream, same step count;
    # the units the kernels process are counted at the start and end of this pass
    def units():
        st = eng.stats()
        touch = None
        if not settling:
            _, fl = eng.friction_state("friction")
            fl = np.asarray(fl)
            touch = int(((fl >> 1) & 1).sum())
        return st, touch
    s1b, t_a = units()
    eng.set_kernel_timing(True)
    t1 = time.perf_counter()
    sim.runTimeSteps(args.steps)
    el_t = time.perf_counter() - t1
    kt = eng.kernel_times()
    eng.set_kernel_timing(False)
    s2, t_b = units()
    log("instrumented %d steps: %.3f s, kernels %s" % (args.steps, el_t, kt))

    nb = 0.5 * (s1b["num_bonds"] + s2["num_bonds"])
    nd = 0.5 * (s1b["num_dynamic"] + s2["num_dynamic"])
    b_part = B_PARTICLE_NROT if settling else B_PARTICLE_ROT
    if settling:
        dyn_bytes = nd * B_ELASTIC
        ntouch = None
    else:
        ntouch = 0.5 * (t_a + t_b)
        dyn_bytes = ntouch * B_FRICTION + (nd - ntouch) * B_FRICTION_IDLE
    per_launch = {"forces": nb * B_BOND + dyn_bytes,
                  "gather_integrate": n * b_part}
    dom = max((k for k in per_launch if kt[k][1] > 0), key=lambda k: kt[k][0])
    avg_s = kt[dom][0] / kt[dom][1] * 1e-3
    achieved = per_launch[dom] / avg_s / 1e9
    flops = {"forces": nb * FLOP_BOND, "gather_integrate": n * FLOP_PARTICLE_GATHER}
    traffic = MEASURED_TRAFFIC.get((args.workload, n), {}).get(dom)
    step_bytes = n * b_part + nb * B_BOND + dyn_bytes
    ms_per_step = el / args.steps * 1e3
    out = {
        "metric": "particle-steps/s",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": ({
            "workload": "frictional shear cell, %d RotSphere, radii U[0.25, 0.75], periodic x, "
                        "RotFriction kn=ks=1e4 mu=0.5, rough bonded walls sliding at +-0.1, "
                        "servo confining stress 1.0, dt=5e-5" % n,
            "particles": n,
            "friction_pairs": int(s1["num_dynamic"]),
            "touching_pairs": None if ntouch is None else int(ntouch),
            "box": [round(b, 3) for b in spec["box"]],
            "steps_per_sync": 1,
        } if shear else {
            "workload": "gravity settling, %d NRotSphere, radii U[0.2, 0.5], density 1, "
                        "NRotElastic k=1, elastic floor k=1, g=9.81, dt=1e-4, box "
                        "20x40x20" % n,
            "particles": n,
            "elastic_pairs": int(s1["num_dynamic"]),
            "box": [20.0, 40.0, 20.0],
            "steps_per_sync": args.batch,
        } if settling else {
            "workload": "bonded rotational uniaxial compression, %d RotSphere, radii "
                        "log-uniform [0.2, 1.0], RotBond kr=1e5 ks=kt=kb=3e4, cohesion "
                        "N(50,5), RotFriction mu=0.6, walls 0.05/time, dt=1e-5" % n,
            "particles": n,
            "bonds": int(s1["num_bonds"]),
            "friction_pairs": int(s1["num_dynamic"]),
            "touching_pairs": None if ntouch is None else int(ntouch),
            "prism": [round(b, 3) for b in spec["box"]],
            "steps_per_sync": args.batch,
            "conditioning_steps": cond,
            "first_timed_step": cond + args.warmup + 1,
        }),
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": None if traffic is None else
            "profiles/step_kernels_bonded100k.md (FETCH_SIZE, WRITE_SIZE of the same command)",
            "algorithmic_bytes_per_launch": per_launch[dom],
            "avg_launch_us": avg_s * 1e6,
        },
        "fp64_roofline": {
            "kernel": dom,
            "flop_per_launch": flops[dom],
            "achieved": flops[dom] / avg_s / 1e9,
            "peak": FP64_PEAK_GFLOPS,
            "unit": "GFLOP/s",
            "frac": flops[dom] / avg_s / 1e9 / FP64_PEAK_GFLOPS,
        } if flops[dom] > 0 else None,
        "step_roofline": {
            "algorithmic_bytes_per_step": step_bytes,
            "achieved_GBs": step_bytes / (ms_per_step * 1e-3) / 1e9,
            "frac": step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        },
        "kernel_ms_per_step": {k: v[0] / max(args.steps, 1) for k, v in kt.items()},
        "host_phases": int(s1["num_host_phases"] - s0["num_host_phases"]),
        "bond_breaks_timed": int(s1["num_broken"] - s0["num_broken"]),
    }
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        last = cond + args.warmup + args.steps  # the last step of the timed pass
        out["cpu_baseline"] = cpu_baseline(pkg, spm, spec, args.cpu_budget, last)[0]
        log("cpu baseline: %.1f s" % (time.perf_counter() - t0))
        import tempfile
        try:
            share = len(os.sched_getaffinity(0))
        except AttributeError:
            share = os.cpu_count() or 1
        workers = args.cpu_workers or max(1, min(share, 16))
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "specimen.pkl")
            save_specimen(path, spec)
            t0 = time.perf_counter()
            out["cpu_baseline_all_cores"] = cpu_baseline_all_cores(path, args.cpu_budget, workers,
                                                                   last)
        log("cpu baseline, %d processes: %.1f s" % (workers, time.perf_counter() - t0))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Particle-steps/s of the plain Brownian-dynamics step at 1M particles on one MI355X.

Workload (plain_1m): 1,000,000 particles in 4 species, 1000 nm periodic box,
harmonic repulsion + piecewise harmonic well (species 0-1), spherical
confinement R = 480 nm, dt = 0.05 ns.  A "step" is one BD step of all
particles: force/update kernel + cell sort, captured in one hipGraph.
Inputs are resident in HBM before the timed region starts.

Prints one JSON line.  --gpus N > 1 is refused: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s at 1M; batched & surrogate vs plain GPU at fixed RDF error"
PEAK_HBM_GBS = 8000.0
B_ALG = 28   # bytes / particle-step: read x,y,z,type (16) + write x,y,z (12), in-kernel noise


def log(msg):
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def read_traffic():
    """HBM bytes per force/update launch from the committed rocprof PMC summary."""
    p = os.path.join(ROOT, "profiles", "force_update_plain_1m_traffic.txt")
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def cpu_baseline(spec, xyz, types, budget_s=15.0):
    """The oracle's FP64 restatement of the reference step (cell list, 27-cell
    scan), single-threaded like Core.step, on a bounded number of steps."""
    from oracle import oracle as orc
    o = orc.OracleSystem(spec)
    n = len(types)
    t0 = time.perf_counter()
    x, _ = o.run(xyz, types, 0, 1, nthreads=1)
    t1 = time.perf_counter() - t0
    steps = max(1, min(50, int(budget_s / max(t1, 1e-3))))
    log(f"cpu baseline: 1 step took {t1:.2f}s, timing {steps} more single-threaded")
    t0 = time.perf_counter()
    o.run(x, types, 1, steps, nthreads=1)
    dt1 = time.perf_counter() - t0
    nth = min(os.cpu_count() or 1, 16)
    t0 = time.perf_counter()
    o.run(x, types, 1, steps, nthreads=nth)
    dtn = time.perf_counter() - t0
    return {
        "value": n * steps / dt1,
        "unit": "particle-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"plain_1m, {steps} steps of all {n} particles, FP64 oracle (oracle/rd_oracle.c orc_run), 1 thread",
        "value_all_cores": n * steps / dtn,
        "cores_all": nth,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cell", type=float, default=0.0, help="lattice cell edge (0: automatic)")
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner")
    ap.add_argument("--kernel", type=int, default=-1, help="force kernel: 1 gather, 2 tiled")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--rows", type=int, default=0)
    ap.add_argument("--fill", type=float, default=0.0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU; --gpus N > 1 is not supported", file=sys.stderr)
        sys.exit(2)

    from readdygpu_amd import Simulation, workloads
    from readdygpu_amd import _lib

    spec, xyz, types = workloads.plain_1m(n=args.n, cell_size=args.cell)
    n = len(types)
    log(f"setting up plain_1m n={n}")
    sim = Simulation(spec, xyz, types)
    if args.kernel >= 0:
        sim.set_launch_config(kernel=args.kernel, block=args.block, rows_per_tile=args.rows, fill=args.fill)
    if not args.no_tune and args.kernel < 0 and args.cell == 0.0:
        ms = sim.autotune(trial_steps=10)
        log(f"autotuned: {ms * 1e3:.1f} us/step")
    info = sim.info()
    lc = sim.launch_config()
    log(f"lattice {list(info.lattice)} cell {info.cell_size:.3f} nm; kernel {lc}; warmup {args.warmup} steps")
    sim.simulate(args.warmup)

    out = (C.c_double * 4)()
    rc = sim.lib.rd_bench_steps(sim.h, args.steps, out)
    if rc != 0:
        raise RuntimeError(sim.lib.rd_last_error(sim.h).decode())
    wall_ms, gpu_ms, force_ms, sort_ms = list(out)
    sim.simulate(0)
    log(f"timed {args.steps} steps: wall {wall_ms:.3f} ms, events {gpu_ms:.3f} ms, "
        f"force/update {force_ms:.3f} ms, sort {sort_ms:.3f} ms")
    ms_per_step = wall_ms / args.steps
    value = n * args.steps / (wall_ms / 1e3)
    force_avg_s = force_ms / args.steps / 1e3
    achieved = B_ALG * n / force_avg_s / 1e9
    roofline = {
        "bound": "hbm",
        "achieved": achieved,
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "frac": achieved / PEAK_HBM_GBS,
        "traffic": read_traffic(),
        "kernel": "k_force_update",
        "kernel_avg_us": force_avg_s * 1e6,
        "bytes_alg_per_launch": B_ALG * n,
        "sort_avg_us": sort_ms / args.steps * 1e3,
    }
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(spec, xyz, types)
    res = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "plain_1m: 1,000,000 particles, 4 species, 1000 nm periodic box, harmonic repulsion k=10 "
                        "+ piecewise well (d=1, l=1) species 0-1, sphere confinement R=480 nm, dt=0.05 ns",
            "n_particles": n,
            "cell_size_nm": info.cell_size,
            "lattice": list(info.lattice),
            "force_kernel": {1: "gather", 2: "tiled", 3: "wave"}.get(lc["used_kernel"], "?"),
            "block": lc["block"],
            "rows_per_tile": lc["used_rows"],
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    sim.close()
    print(json.dumps(res))


if __name__ == "__main__":
    main()

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
import dataclasses
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
        "host": host_cpu(),
    }


def host_cpu():
    """CPU model and core counts of the host the baseline ran on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = None
    return {"model": model, "nproc": os.cpu_count(), "usable": usable}


def tuned_config(sim):
    """(cell size, launch config) the handle runs with, for another handle of
    the same workload to adopt"""
    lc = sim.launch_config()
    return sim.info().cell_size, {"kernel": lc["used_kernel"], "block": lc["used_block"], "ppt": lc["used_ppt"]}


def adopt_config(sim, cfg):
    cell, lc = cfg
    sim.set_cell_size(cell)
    sim.set_launch_config(**lc)


def bench_handle(sim, steps, warmup, tune_steps=0, direct=False):
    """(wall ms of the graph pass, force ms, sort ms, external-pass ms of the
    event pass) for `steps` steps after `warmup` steps; with tune_steps the
    autotuner runs on the warmed-up state (it restores the state, so the
    trajectory does not change).  direct: no graph capture (the wall time is
    then that of the directly launched, event-timed steps)."""
    sim.simulate(warmup)
    if tune_steps:
        sim.autotune(trial_steps=tune_steps)
    out = (C.c_double * 5)()
    fn = sim.lib.rd_bench_steps_direct if direct else sim.lib.rd_bench_steps
    rc = fn(sim.h, steps, out)
    if rc != 0:
        raise RuntimeError(sim.lib.rd_last_error(sim.h).decode())
    sim.simulate(0)
    return out[0], out[2], out[3], out[4]


def fp64_error(steps=20):
    """FP32 GPU trajectory against the FP64 oracle on the 10k workload (same
    start, same noise): the largest minimum-image position difference after
    one step and after `steps` steps of a free run."""
    import numpy as np
    from oracle import oracle as orc
    from readdygpu_amd import Simulation, workloads
    spec, xyz, types = workloads.plain_10k()
    o = orc.OracleSystem(spec)
    L = np.asarray(spec.box_hi, float) - np.asarray(spec.box_lo, float)

    def err(a, b):
        d = np.abs(np.asarray(a, float) - np.asarray(b, float))
        return float(np.minimum(d, L - d).max())

    with Simulation(spec, xyz, types) as sim:
        sim.simulate(1)
        x1, _ = sim.copy_pos_from_device()
        sim.simulate(steps - 1)
        xk, _ = sim.copy_pos_from_device()
    r1, b1 = o.run(xyz, types, 0, 1)
    rk, bk = o.run(xyz, types, 0, steps)
    res = {"workload": "plain_10k: 10,000 particles, 2 species, 100 nm box, harmonic repulsion, dt=0.1 ns; "
                       "GPU FP32 vs oracle FP64 (oracle/rd_oracle.c orc_run), same start and noise",
           "one_step_max_nm": err(x1, r1), "steps": steps, "free_run_max_nm": err(xk, rk),
           "oracle_flags": int(b1 | bk)}
    log(f"fp64 error: one step {res['one_step_max_nm']:.2e} nm, {steps} steps {res['free_run_max_nm']:.2e} nm")
    return res


def configured_run(spec, xyz, types, lc):
    """The 1M workload as configured (500 steps from the initial state) with
    the tuned cell size and force-kernel form and shape: host wall time of the 500 steps as
    one graph (rd_bench_steps pass 1), and the force-kernel average over the
    500 steps after them."""
    from readdygpu_amd import Simulation, workloads
    steps = workloads.PLAIN_1M["steps"]
    with Simulation(spec, xyz, types) as sim:
        sim.set_launch_config(kernel=lc["used_kernel"], block=lc["used_block"], ppt=lc["used_ppt"])
        out = (C.c_double * 5)()
        rc = sim.lib.rd_bench_steps(sim.h, steps, out)
        if rc != 0:
            raise RuntimeError(sim.lib.rd_last_error(sim.h).decode())
    n = len(types)
    res = {"steps": steps, "value": n * steps / (out[0] / 1e3), "ms_per_step": out[0] / steps,
           "window": f"steps 0-{steps} from the initial state (one graph, host clock)",
           "force_avg_us_steps_500_1000": out[2] / steps * 1e3, "sort_avg_us_steps_500_1000": out[3] / steps * 1e3}
    log(f"configured run: {steps} steps {res['ms_per_step'] * 1e3:.1f} us/step ({res['value']:.3e} p-steps/s)")
    return res


def bench_batched(warmup=10, plain_steps=100):
    """256 replicas (log-uniform 500..50,000 particles, phi = 0.05) in one
    handle over the configured 1,000 steps from the initial state, in both
    force-kernel forms (one workgroup per group of work items, and the
    resident grid whose workgroups take their share of items from LDS tickets) at the tuned cell
    and workgroup size, against the same replicas run one handle at a time."""
    from readdygpu_amd import Simulation, workloads
    steps = workloads.BATCHED["steps"]
    spec, boxes, counts, xyz, types = workloads.batched_replicas()
    n = int(counts.sum())
    sim = Simulation(spec, xyz, types, replicas=(boxes, counts))
    sim.simulate(warmup)
    ms_tune = sim.autotune(trial_steps=5)
    info = sim.info()
    lc = sim.launch_config()
    sim.close()
    forms = {}
    for name, kern in (("per_workgroup", 1), ("work_queue", 2)):
        with Simulation(dataclasses.replace(spec, cell_size=info.cell_size), xyz, types,
                        replicas=(boxes, counts)) as one:
            one.set_launch_config(kernel=kern, block=lc["used_block"], ppt=lc["used_ppt"] if kern == 1 else 1)
            out = (C.c_double * 5)()
            rc = one.lib.rd_bench_steps(one.h, steps, out)
            if rc != 0:
                raise RuntimeError(one.lib.rd_last_error(one.h).decode())
        forms[name] = {"value": n * steps / (out[0] / 1e3), "ms_per_step": out[0] / steps,
                       "force_avg_us_next_1000": out[2] / steps * 1e3, "sort_avg_us_next_1000": out[3] / steps * 1e3}
        log(f"batched {name}: {forms[name]['value']:.3e} p-steps/s ({forms[name]['ms_per_step'] * 1e3:.1f} us/step, "
            f"force {forms[name]['force_avg_us_next_1000']:.1f} us)")
    best = max(forms, key=lambda k: forms[k]["value"])
    batched = forms[best]["value"]
    # the plain form: each replica on its own handle, all steps of one before the next
    tot_ms = 0.0
    off = 0
    for r, c in enumerate(counts):
        b = boxes[r]
        s1 = dataclasses.replace(spec, box_lo=(b[0], b[2], b[4]), box_hi=(b[1], b[3], b[5]), stream=r)
        one = Simulation(s1, xyz[off:off + c], types[off:off + c])
        w1, _, _, _ = bench_handle(one, plain_steps, 2)
        one.close()
        tot_ms += w1
        off += c
    separate = n * plain_steps / (tot_ms / 1e3)
    log(f"batched: {batched:.3e} p-steps/s ({best}), one handle per replica {separate:.3e}")
    return {
        "workload": "batched_replicas: 256 periodic boxes, N log-uniform in [500, 50000] (seed 3), phi=0.05, "
                    "2 species, harmonic repulsion, dt=0.1 ns",
        "n_particles": n, "n_replicas": int(len(counts)), "steps": steps,
        "window": f"steps 0-{steps} from the initial state, one graph, host clock",
        "value": batched, "unit": "particle-steps/s", "form": best, "forms": forms,
        "plain_per_replica_value": separate, "plain_per_replica_steps": plain_steps,
        "speedup_vs_plain": batched / separate,
        "cell_size_nm": info.cell_size, "workgroup": lc["used_block"], "particles_per_lane": lc["used_ppt"],
        "tuned_kernel": lc["used_kernel"], "autotune_us_per_step": ms_tune * 1e3,
    }


def rdf_l2(g_a, g_b):
    """relative L2 deviation of g(r) curves (all type pairs and bins)"""
    import numpy as np
    return float(np.linalg.norm(g_a - g_b) / max(np.linalg.norm(g_b), 1e-30))


def bench_learned(steps=100, warmup=10, train_steps=3000, samples=4_000_000, direct=False, save=None):
    """2M particles in a 2000x2000x400 nm slab with the 512-piece composite:
    exact external pass against the trained BF16 surrogate."""
    import numpy as np
    from readdygpu_amd import Simulation, workloads
    spec, pieces, xyz, types = workloads.learned_slab()
    n = len(types)
    res = {"workload": "learned_slab: 2,000,000 particles, 3 species, 2000x2000x400 nm, harmonic repulsion, "
                       "512-piece nearest-piece composite (k=5, radii 10-80 nm), dt=0.05 ns",
           "n_particles": n, "steps": steps}
    sim = Simulation(spec, xyz, types, rdf=(20.0, 200))
    sim.set_composite(pieces)
    t0 = time.perf_counter()
    sim.surrogate_init(seed=4)
    gs = sim.surrogate_make_data(samples, seed=4)
    t1 = time.perf_counter()
    curve = sim.surrogate_train(steps=train_steps, batch=8192, lr=3e-3, lr_min=1e-5, seed=4)
    t2 = time.perf_counter()
    rmse = sim.surrogate_validate(400_000, seed=12345)
    res.update({"train_samples": samples, "train_steps": train_steps, "train_batch": 8192,
                "data_s": t1 - t0, "train_s": t2 - t1, "loss_first": float(curve[0]), "loss_last": float(curve[-1]),
                "force_rel_rmse": rmse, "gscale": gs})
    log(f"learned: data {t1 - t0:.2f}s train {t2 - t1:.2f}s loss {curve[0]:.4f}->{curve[-1]:.5f} "
        f"rel force RMSE {rmse:.4f}")
    w = sim.surrogate_weights()
    lo, hi, gsc = sim.surrogate_norm()
    if save:
        np.savez(save, w=w, lo=lo, hi=hi, gscale=gsc)
    # exact external pass (this handle), then a fresh handle with the same
    # start, seed and step counter running the surrogate: identical noise; the
    # surrogate handle runs the lattice and launch shape tuned for the exact
    # one, so that only the external pass differs
    wall_e, force_e, sort_e, ext_e = bench_handle(sim, steps, warmup, tune_steps=10, direct=direct)
    cfg = tuned_config(sim)
    sim.call_rdf_calculation()
    g_exact = sim.normalize_rdf()
    sim.close()
    sim = Simulation(spec, xyz, types, rdf=(20.0, 200))
    sim.set_composite(pieces)
    sim.set_surrogate_weights(w)
    sim.set_surrogate_norm(lo, hi, gsc)
    sim.set_external_mode(True)
    adopt_config(sim, cfg)
    wall_s, force_s, sort_s, ext_s = bench_handle(sim, steps, warmup, direct=direct)
    sim.call_rdf_calculation()
    g_sur = sim.normalize_rdf()
    sim.close()
    outs = {}
    for name, (w, f, so, e) in {"exact": (wall_e, force_e, sort_e, ext_e),
                                "surrogate": (wall_s, force_s, sort_s, ext_s)}.items():
        outs[name] = {"value": n * steps / (w / 1e3), "ms_per_step": w / steps, "ext_avg_us": e / steps * 1e3,
                      "force_avg_us": f / steps * 1e3, "sort_avg_us": so / steps * 1e3}
    f_alg = 17152.0
    outs["surrogate"]["mfma"] = {"flop_alg_per_particle": f_alg,
                                 "achieved_tflops": f_alg * n / (ext_s / steps / 1e3) / 1e12,
                                 "peak_tflops": 2500.0}
    outs["surrogate"]["mfma"]["frac"] = outs["surrogate"]["mfma"]["achieved_tflops"] / 2500.0
    res.update(outs)
    res["summed"] = summed_slab(spec, pieces, xyz, types, samples, train_steps)
    res["speedup_surrogate_vs_exact_step"] = outs["surrogate"]["value"] / outs["exact"]["value"]
    res["speedup_surrogate_vs_exact_external"] = ext_e / max(ext_s, 1e-12)
    res["rdf_rel_l2"] = rdf_l2(g_sur, g_exact)
    log(f"learned: exact {outs['exact']['value']:.3e} p-steps/s (ext {outs['exact']['ext_avg_us']:.0f} us), "
        f"surrogate {outs['surrogate']['value']:.3e} (ext {outs['surrogate']['ext_avg_us']:.0f} us), "
        f"RDF rel L2 {res['rdf_rel_l2']:.4f}")
    return res


def learned_kernels(weights, steps=20):
    """The learned slab's external passes and force kernel for kernel-trace
    profiling: no training and no graph capture in this process (the trained
    weights come from a file written by --only learned_slab --save-surrogate);
    `steps` steps each with the exact pass and with the surrogate, launched
    directly."""
    import numpy as np
    from readdygpu_amd import Simulation, workloads
    spec, pieces, xyz, types = workloads.learned_slab()
    z = np.load(weights)
    res = {}
    for name, sur in (("exact", False), ("surrogate", True)):
        with Simulation(spec, xyz, types) as sim:
            sim.set_composite(pieces)
            if sur:
                sim.set_surrogate_weights(z["w"])
                sim.set_surrogate_norm(z["lo"], z["hi"], float(z["gscale"]))
                sim.set_external_mode(True)
            out = (C.c_double * 5)()
            rc = sim.lib.rd_bench_steps_direct(sim.h, steps, out)
            if rc != 0:
                raise RuntimeError(sim.lib.rd_last_error(sim.h).decode())
        res[name] = {"ext_avg_us": out[4] / steps * 1e3, "force_avg_us": out[2] / steps * 1e3,
                     "sort_avg_us": out[3] / steps * 1e3}
        log(f"learned kernels ({name}): ext {res[name]['ext_avg_us']:.1f} us, force {res[name]['force_avg_us']:.1f} us")
    return res


def ext_pass_ms(sim, reps=20):
    ms = C.c_double()
    rc = sim.lib.rd_bench_ext(sim.h, reps, C.byref(ms))
    if rc != 0:
        raise RuntimeError(sim.lib.rd_last_error(sim.h).decode())
    return ms.value


def summed_slab(spec, pieces, xyz, types, samples, train_steps):
    """The 512 pieces summed, as orderOne sums every order-one potential: the
    surrogate's accuracy on it and both external passes on the initial 2M
    positions.  The step itself diverges at this dt and these D (the summed
    curvature, d2U/dz2 >= 1280 kJ/mol/nm^2, puts dt D K / kT at 6.6-26, above
    the Euler-Maruyama bound of 2), so no trajectory or RDF is taken."""
    from readdygpu_amd import Simulation
    n = len(types)
    with Simulation(spec, xyz, types) as sim:
        sim.set_composite(pieces, combine="sum")
        ms_exact = ext_pass_ms(sim)
        sim.surrogate_init(seed=4)
        sim.surrogate_make_data(samples, seed=4)
        curve = sim.surrogate_train(steps=train_steps, batch=8192, lr=3e-3, lr_min=1e-5, seed=4)
        rmse = sim.surrogate_validate(400_000, seed=12345)
        sim.set_external_mode(True)
        ms_sur = ext_pass_ms(sim)
    f_alg = 17152.0
    res = {"force_rel_rmse": rmse, "loss_last": float(curve[-1]), "exact_ext_us": ms_exact * 1e3,
           "surrogate_ext_us": ms_sur * 1e3, "speedup_surrogate_vs_exact_external": ms_exact / ms_sur,
           "mfma_tflops": f_alg * n / (ms_sur / 1e3) / 1e12,
           "note": "external passes alone on the initial positions; the summed potential's step diverges"}
    log(f"learned (summed pieces): rel force RMSE {rmse:.5f}, exact pass {ms_exact * 1e3:.0f} us, "
        f"surrogate pass {ms_sur * 1e3:.0f} us")
    return res


def bench_capacity(samples=4_000_000):
    """Where the 3-64-64-64-3 surrogate's accuracy runs out, on the learned
    slab's box (nearest-piece composites of its first P pieces): relative force
    RMSE on fresh samples after 3,000 optimiser steps for P = 1 ... 512, and
    for P = 512 after 10,000 and 30,000 steps (more training does not close
    the gap: the error is set by the network's capacity against the number of
    switch surfaces between pieces)."""
    from readdygpu_amd import Simulation, workloads
    spec, pieces, xyz, types = workloads.learned_slab(n=20_000)
    res = {"workload": "learned_slab box (2000x2000x400 nm), nearest-piece composite of the first P of its 512 "
                       "pieces; 3-64-64-64-3 BF16 MLP, batch 8192, lr 3e-3 cosine to 1e-5, 4M samples",
           "pieces": {}, "training_steps_512": {}}
    with Simulation(spec, xyz, types) as sim:
        for p, steps in [(1, 3000), (4, 3000), (16, 3000), (64, 3000), (512, 3000), (512, 10000), (512, 30000)]:
            sim.set_composite(pieces[:p])
            sim.surrogate_init(seed=4)
            sim.surrogate_make_data(samples, seed=4)
            t0 = time.perf_counter()
            sim.surrogate_train(steps=steps, batch=8192, lr=3e-3, lr_min=1e-5, seed=4)
            t1 = time.perf_counter()
            rmse = sim.surrogate_validate(400_000, seed=12345)
            if steps == 3000:
                res["pieces"][str(p)] = rmse
            if p == 512:
                res["training_steps_512"][str(steps)] = rmse
            log(f"capacity: {p} pieces, {steps} steps ({t1 - t0:.2f} s): rel force RMSE {rmse:.4f}")
    return res


def bench_membrane(steps=200, frames=20):
    """The accuracy gates where the geometry is learnable: a membrane layer
    (one disk wider than a 400 x 400 x 40 nm box, 120,000 particles).
    Surrogate vs exact from the same start with the same noise."""
    import numpy as np
    from readdygpu_amd import Simulation, SystemSpec
    from readdygpu_amd.model import (ATTRACTIVE, DISK, HARMONIC, REPULSIVE, OrderOnePotential, OrderTwoPotential,
                                     ParticleType)
    from readdygpu_amd import workloads
    spec = SystemSpec(box_lo=(0.0, 0.0, 0.0), box_hi=(400.0, 400.0, 40.0), dt=0.05, seed=6)
    spec.types = [ParticleType("A", D=1.0, radius=1.0), ParticleType("B", D=0.5, radius=1.5)]
    spec.pot2 = [OrderTwoPotential(HARMONIC, REPULSIVE, k=10.0)]
    pieces = [OrderOnePotential(DISK, ATTRACTIVE, k=5.0, origin=(200.0, 200.0, 20.0), normal=(0.0, 0.0, 1.0),
                                radius=800.0)]
    rng = np.random.default_rng(6)
    n = 120_000
    xyz = workloads.uniform_positions(rng, n, spec.box_lo, spec.box_hi)
    types = (np.arange(n) % 2).astype(np.int32)
    g = {}
    w = lo = hi = gs = rmse = None
    for mode in ("exact", "surrogate"):
        sim = Simulation(spec, xyz, types, rdf=(10.0, 50))
        sim.set_composite(pieces)
        if mode == "exact":
            sim.surrogate_init(seed=4)
            sim.surrogate_make_data(1 << 21, seed=4)
            sim.surrogate_train(steps=3000, batch=8192, lr=3e-3, lr_min=1e-5)
            rmse = sim.surrogate_validate(200_000, seed=99)
            w = sim.surrogate_weights()
            lo, hi, gs = sim.surrogate_norm()
        else:
            sim.set_surrogate_weights(w)
            sim.set_surrogate_norm(lo, hi, gs)
            sim.set_external_mode(True)
        sim.simulate(steps)
        for _ in range(frames):
            sim.simulate(10)
            sim.call_rdf_calculation()
        g[mode] = sim.normalize_rdf()
        sim.close()
    dev = rdf_l2(g["surrogate"], g["exact"])
    log(f"membrane: force rel RMSE {rmse:.4f}, RDF rel L2 {dev:.4f}")
    return {"workload": "membrane layer: 120,000 particles, 400x400x40 nm, one disk wider than the box (k=5), "
                        "surrogate vs exact from the same start and noise",
            "force_rel_rmse": rmse, "rdf_rel_l2": dev, "gates": {"force_rel_rmse": 0.01, "rdf_rel_l2": 0.005},
            "passed": bool(rmse <= 0.01 and dev <= 0.005)}


def bench_passes(sim, reps=10):
    """The per-invocation passes beside the step on the same state: RDF
    (0-20 nm, 200 bins, all type pairs), A+B reaction candidates at 4 nm,
    and the neighbour-pair emission."""
    sim.configure_rdf(20.0, 200)
    sim.call_rdf_calculation()
    sim.copy_rdf_matrix()
    t0 = time.perf_counter()
    for _ in range(reps):
        sim.call_rdf_calculation()
    sim.copy_rdf_matrix()
    t_rdf = (time.perf_counter() - t0) / reps
    c = sim.reaction_candidates()
    t0 = time.perf_counter()
    for _ in range(reps):
        c = sim.reaction_candidates()
    t_cand = (time.perf_counter() - t0) / reps
    p = sim.neighbor_pairs()
    t0 = time.perf_counter()
    for _ in range(reps):
        p = sim.neighbor_pairs()
    t_pairs = (time.perf_counter() - t0) / reps
    res = {"rdf_ms": t_rdf * 1e3, "rdf": "0-20 nm, 200 bins, 4x4 type pairs",
           "candidates_ms": t_cand * 1e3, "candidates": int(len(c)),
           "neighbor_pairs_ms": t_pairs * 1e3, "neighbor_pairs": int(len(p)),
           "note": "host wall time per call, including the copy of the pair lists to the host"}
    log(f"passes: RDF {res['rdf_ms']:.2f} ms, candidates {res['candidates_ms']:.2f} ms ({len(c)}), "
        f"neighbour pairs {res['neighbor_pairs_ms']:.2f} ms ({len(p)})")
    return res


def bench_frames(frames=20, spf=100):
    """Frame hand-off on the 1M workload: the sequential loop of the
    reference (simulate, synchronous copy, host packs JPos = [N, (x,y,z,
    index,id) x N]) against the loop that runs frame f+1 during the hand-off
    of frame f (pinned asynchronous copy); and the device-side application of
    a cReactions list (100 removals + 100 creations)."""
    import numpy as np
    from readdygpu_amd import Simulation, workloads
    spec, xyz, types = workloads.plain_1m()
    n = len(types)
    ids = np.arange(n, dtype=np.float32)

    def jpos(x):
        a = np.empty((len(x), 5), np.float32)
        a[:, :3] = x
        a[:, 3] = ids[:len(x)]
        a[:, 4] = ids[:len(x)]
        return a

    res = {"workload": f"plain_1m, {frames} frames x {spf} steps, JPos packed on the host per frame"}
    with Simulation(spec, xyz, types) as sim:
        sim.autotune(trial_steps=2)
        sim.simulate(spf)
        t0 = time.perf_counter()
        sim.simulate(frames * spf)
        t_gpu = time.perf_counter() - t0
        t0 = time.perf_counter()
        for _ in range(frames):
            sim.simulate(spf)
            x, _ = sim.copy_pos_from_device()
            jpos(x)
        t_seq = time.perf_counter() - t0
        t0 = time.perf_counter()
        sim.simulate_async(spf)
        for f in range(frames):
            sim.frame_copy()
            if f + 1 < frames:
                sim.simulate_async(spf)
            x, _ = sim.frame_wait(copy=False)
            jpos(x)
        t_ovl = time.perf_counter() - t0
        rng = np.random.default_rng(1)
        ts = []
        for _ in range(5):
            m = sim.info().n
            dele = rng.choice(m, 100, replace=False)
            e = [200]
            for i in dele:
                e += [0, -1, 0, 0, 0, int(i)]
            for k in range(100):
                e += [0, k % 4, float(rng.uniform(1, 999)), float(rng.uniform(1, 999)), float(rng.uniform(1, 999)), -1]
            cr = np.array(e, np.float32)
            t1 = time.perf_counter()
            sim.apply_reactions(cr)
            ts.append(time.perf_counter() - t1)
    ps = n * spf * frames
    res.update({"gpu_only_value": ps / t_gpu, "sequential_value": ps / t_seq, "overlapped_value": ps / t_ovl,
                "unit": "particle-steps/s", "handoff_ms_per_frame_sequential": (t_seq - t_gpu) / frames * 1e3,
                "handoff_ms_per_frame_overlapped": (t_ovl - t_gpu) / frames * 1e3,
                "apply_reactions_ms": float(np.median(ts)) * 1e3})
    log(f"frames: gpu only {res['gpu_only_value']:.3e}, sequential {res['sequential_value']:.3e}, overlapped "
        f"{res['overlapped_value']:.3e} p-steps/s; reactions {res['apply_reactions_ms']:.2f} ms")
    return res


def cytosol_reactions(spec, boxes, counts, xyz, types, pieces, mtypes, w, norm, frames=10, spf=100, p=0.2,
                      c_type=2):
    """A + B -> C at 3 nm in the batched cytosol with the surrogate membrane:
    every spf steps the candidate pairs (species 0 and 1 within 3 nm) react
    with probability p; a particle reacts at most once per frame (an accepted
    pair reacts when its random priority is the lowest at both particles); A
    is removed and B becomes C (species c_type) at the pair's midpoint, in its
    replica."""
    import numpy as np
    from readdygpu_amd import Simulation
    rng = np.random.default_rng(55)
    cnt = np.array(counts, dtype=np.int64)
    sim = Simulation(spec, xyz, types, replicas=(boxes, counts))
    sim.set_composite(pieces, types=mtypes)
    sim.set_surrogate_weights(w)
    sim.set_surrogate_norm(*norm)
    sim.set_external_mode(True)
    sim.simulate(10)
    nreact, t_cand, t_dec, t_apply = 0, 0.0, 0.0, 0.0
    lo_all, hi_all = boxes[:, 0::2], boxes[:, 1::2]
    t0 = time.perf_counter()
    for f in range(frames):
        sim.simulate(spf)
        t1 = time.perf_counter()
        c = sim.reaction_candidates()
        sim.frame_copy()
        x, t = sim.frame_wait(copy=False)
        t2 = time.perf_counter()
        acc = c[rng.random(len(c)) < p]
        pri = rng.random(len(acc))
        best = np.full(len(t), np.inf)
        np.minimum.at(best, acc[:, 0], pri)
        np.minimum.at(best, acc[:, 1], pri)
        sel = acc[(best[acc[:, 0]] == pri) & (best[acc[:, 1]] == pri)]
        a = np.where(t[sel[:, 0]] == 0, sel[:, 0], sel[:, 1])
        b = np.where(t[sel[:, 0]] == 0, sel[:, 1], sel[:, 0])
        offs = np.concatenate([[0], np.cumsum(cnt)[:-1]])
        r = np.searchsorted(offs, a, side="right") - 1
        lo, hi = lo_all[r], hi_all[r]
        L = hi - lo
        d = x[b].astype(np.float64) - x[a]
        d -= L * np.round(d / L)
        mid = np.minimum(np.mod(x[a] + 0.5 * d - lo, L) + lo, hi - 1e-3)
        k = len(sel)
        ent = np.zeros((k, 2, 6), np.float32)
        ent[:, 0, 1] = -1
        ent[:, 0, 5] = a
        ent[:, 1, 1] = c_type
        ent[:, 1, 2:5] = mid
        ent[:, 1, 5] = b
        np.subtract.at(cnt, r, 1)
        t3 = time.perf_counter()
        if k:
            sim.apply_reactions(np.concatenate([[2 * k], ent.reshape(-1)]).astype(np.float32),
                                creation_replica=np.repeat(r, 2).astype(np.int32))
        t4 = time.perf_counter()
        nreact += k
        t_cand += t2 - t1
        t_dec += t3 - t2
        t_apply += t4 - t3
    wall = time.perf_counter() - t0
    info = sim.info()
    sim.close()
    ps = float(np.sum(counts)) * frames * spf
    out = {"frames": frames, "steps_per_frame": spf, "reactions": nreact, "n_after": info.n,
           "value_incl_reactions": ps / wall, "candidates_and_copy_ms_per_frame": t_cand / frames * 1e3,
           "host_decision_ms_per_frame": t_dec / frames * 1e3, "apply_ms_per_frame": t_apply / frames * 1e3}
    log(f"cytosol reactive: {nreact} A+B->C in {frames} frames, {out['value_incl_reactions']:.3e} p-steps/s incl. "
        f"reactions; per frame: candidates + copy {out['candidates_and_copy_ms_per_frame']:.1f} ms, host decision "
        f"{out['host_decision_ms_per_frame']:.1f} ms, device compaction {out['apply_ms_per_frame']:.1f} ms")
    return out


def cytosol_rdf(sim, every):
    """g(r) per replica over the configured 1,000 steps, a frame every `every` steps."""
    from readdygpu_amd import workloads
    steps = workloads.CYTOSOL["steps"]
    for _ in range(steps // every):
        sim.simulate(every)
        sim.call_rdf_calculation()
    g = sim.normalize_rdf()
    sim.close()
    return g


def bench_cytosol(steps=100, warmup=10, rdf_every=2, plain_steps=20, train_steps=3000):
    """Combined workload: 64 replicas (5.06M particles, phi = 0.3, 8 species,
    two wells, A+B candidates) with a 256-disk membrane acting on the two
    largest species.  Batched + surrogate against batched exact (the same
    replicas, same start and noise: bitwise what each replica gives alone)
    and against the plain form (one handle per replica, exact membrane)."""
    import dataclasses
    import numpy as np
    from readdygpu_amd import Simulation, workloads
    spec, boxes, counts, xyz, types, pieces, mtypes = workloads.crowded_cytosol()
    n = int(counts.sum())
    lmax = float(boxes[:, 1].max())
    res = {"workload": "crowded_cytosol: 64 replicas, N log-uniform in [20000, 200000] (seed 5), phi=0.3, 8 species, "
                       "2 wells, A+B candidates at 3 nm, 256-disk membrane on species 6-7, dt=0.01 ns",
           "n_particles": n, "n_replicas": int(len(counts)), "steps": steps}

    def handle(surrogate, w=None, norm=None, sp=spec):
        sim = Simulation(sp, xyz, types, replicas=(boxes, counts), rdf=(20.0, 100))
        sim.set_composite(pieces, types=mtypes)
        if surrogate:
            sim.set_surrogate_weights(w)
            sim.set_surrogate_norm(*norm)
            sim.set_external_mode(True)
        return sim

    sim = handle(False)
    sim.surrogate_init(seed=5)
    sim.set_surrogate_norm((0.0, 0.0, 0.0), (lmax, lmax, lmax), 1.0)
    sim.surrogate_make_data(4_000_000, seed=5)
    sim.surrogate_train(steps=train_steps, batch=8192, lr=3e-3, lr_min=1e-5, seed=5)
    res["force_rel_rmse"] = sim.surrogate_validate(400_000, seed=55)
    w = sim.surrogate_weights()
    norm = sim.surrogate_norm()
    out = {}
    g = {}
    cfg = None
    for mode in ("exact", "surrogate"):
        if mode == "surrogate":
            # the lattice and launch shape tuned for the exact handle
            sim = handle(True, w, norm)
            adopt_config(sim, cfg)
        wall, force, sort, ext = bench_handle(sim, steps, warmup, tune_steps=2 if cfg is None else 0)
        if cfg is None:
            cfg = tuned_config(sim)
        out[mode] = {"value": n * steps / (wall / 1e3), "ms_per_step": wall / steps,
                     "ext_avg_us": ext / steps * 1e3, "force_avg_us": force / steps * 1e3,
                     "sort_avg_us": sort / steps * 1e3}
        out[mode]["candidates_ab"] = int(len(sim.reaction_candidates()))
        sim.close()
        # RDF over the configured run from the common start (fresh handle, same noise)
        g[mode] = cytosol_rdf(handle(mode == "surrogate", w, norm), rdf_every)
    # statistical floor: the exact run with another noise seed
    g["exact_seed2"] = cytosol_rdf(handle(False, sp=dataclasses.replace(spec, seed=spec.seed + 1000)), rdf_every)

    def rel(a, b):
        per = [float(np.linalg.norm(g[a][r] - g[b][r]) / max(np.linalg.norm(g[b][r]), 1e-30))
               for r in range(len(counts))]
        pa = (g[a] * counts[:, None, None, None]).sum(0)
        pb = (g[b] * counts[:, None, None, None]).sum(0)
        return {"mean": float(np.mean(per)), "max": float(np.max(per)),
                "pooled": float(np.linalg.norm(pa - pb) / np.linalg.norm(pb))}
    res.update(out)
    steps_cfg = workloads.CYTOSOL["steps"]
    res["rdf_sampling"] = (f"{steps_cfg} steps from the initial state, a frame every {rdf_every} steps "
                           f"({steps_cfg // rdf_every} frames), 0-20 nm in 100 bins, relative L2 per replica")
    res["rdf_rel_l2_surrogate_vs_exact"] = rel("surrogate", "exact")
    res["rdf_rel_l2_exact_other_seed_vs_exact"] = rel("exact_seed2", "exact")
    res["rdf_gate"] = 0.005
    # plain form: each replica on its own handle with the exact membrane
    tot = 0.0
    off = 0
    for r, c in enumerate(counts):
        b = boxes[r]
        s1 = dataclasses.replace(spec, box_lo=(b[0], b[2], b[4]), box_hi=(b[1], b[3], b[5]), stream=r)
        one = Simulation(s1, xyz[off:off + c], types[off:off + c])
        one.set_composite(pieces, types=mtypes)
        w1, _, _, _ = bench_handle(one, plain_steps, 2)
        one.close()
        tot += w1
        off += c
    res["plain_exact_per_replica_value"] = n * plain_steps / (tot / 1e3)
    res["reactive"] = cytosol_reactions(spec, boxes, counts, xyz, types, pieces, mtypes, w, norm)
    res["speedup_batched_surrogate_vs_plain"] = out["surrogate"]["value"] / res["plain_exact_per_replica_value"]
    res["speedup_batched_surrogate_vs_batched_exact"] = out["surrogate"]["value"] / out["exact"]["value"]
    log(f"cytosol: batched exact {out['exact']['value']:.3e}, batched surrogate {out['surrogate']['value']:.3e}, "
        f"plain exact {res['plain_exact_per_replica_value']:.3e} p-steps/s; force RMSE {res['force_rel_rmse']:.4f}; "
        f"RDF rel L2 surrogate {res['rdf_rel_l2_surrogate_vs_exact']}, other seed {res['rdf_rel_l2_exact_other_seed_vs_exact']}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cell", type=float, default=0.0, help="lattice cell edge (0: automatic)")
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner")
    ap.add_argument("--kernel", type=int, default=-1, help="force kernel: 1 gather")
    ap.add_argument("--block", type=int, default=0, help="force-kernel workgroup size (64-512; 0: autotuned)")
    ap.add_argument("--ppt", type=int, default=0, help="force-kernel particles per lane (1, 2, 4; 0: autotuned)")
    ap.add_argument("--no-extras", action="store_true", help="skip the batched and learned workloads")
    ap.add_argument("--only", choices=["batched_replicas", "learned_slab", "learned_membrane", "crowded_cytosol",
                                       "frame_handoff", "learned_capacity"],
                    default=None,
                    help="run only this extra workload and print its JSON object")
    ap.add_argument("--direct", action="store_true",
                    help="with --only learned_slab: time the steps without graph capture (for profilers)")
    ap.add_argument("--save-surrogate", default=None, metavar="NPZ",
                    help="with --only learned_slab: write the trained weights and normalisation")
    ap.add_argument("--learned-kernels", default=None, metavar="NPZ",
                    help="only launch the learned slab's kernels with the weights of NPZ, no training and no graph "
                         "capture (kernel-trace profiling)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel over BYTES bytes (profiler HBM-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU; --gpus N > 1 is not supported", file=sys.stderr)
        sys.exit(2)

    from readdygpu_amd import Simulation, workloads

    if args.learned_kernels:
        print(json.dumps(learned_kernels(args.learned_kernels)))
        return
    if args.only:
        fn = {"batched_replicas": bench_batched, "learned_slab": bench_learned,
              "learned_membrane": bench_membrane, "crowded_cytosol": bench_cytosol,
              "frame_handoff": bench_frames, "learned_capacity": bench_capacity}[args.only]
        if args.only == "learned_slab":
            print(json.dumps(fn(direct=args.direct, save=args.save_surrogate)))
        else:
            print(json.dumps(fn()))
        return
    spec, xyz, types = workloads.plain_1m(n=args.n, cell_size=args.cell)
    if args.calibrate_copy:
        sim = Simulation(spec, xyz[:1000], types[:1000])
        ms = C.c_double()
        rc = sim.lib.rd_bench_copy(sim.h, int(args.calibrate_copy), 10, C.byref(ms))
        if rc != 0:
            raise RuntimeError(sim.lib.rd_last_error(sim.h).decode())
        sim.close()
        print(json.dumps({"copy_bytes": args.calibrate_copy, "reps": 10, "avg_ms": ms.value,
                          "gbs": 2 * args.calibrate_copy / (ms.value / 1e3) / 1e9}))
        return
    n = len(types)
    log(f"setting up plain_1m n={n}")
    sim = Simulation(spec, xyz, types)
    if args.kernel >= 0 or args.block or args.ppt:
        sim.set_launch_config(kernel=max(args.kernel, 0), block=args.block, ppt=args.ppt)
    # warm up first: the confinement pulls the particles outside the sphere
    # onto its surface within the first steps, so the tuned cell size has to
    # be chosen on the shell that the timed steps see (the autotuner restores
    # the state it starts from)
    sim.simulate(args.warmup)
    if not args.no_tune and args.kernel < 0 and args.cell == 0.0 and not args.block and not args.ppt:
        ms = sim.autotune(trial_steps=20)
        log(f"autotuned: {ms * 1e3:.1f} us/step")
    info = sim.info()
    lc = sim.launch_config()
    log(f"lattice {list(info.lattice)} cell {info.cell_size:.3f} nm; kernel {lc}; after {args.warmup} warmup steps")

    out = (C.c_double * 5)()
    rc = sim.lib.rd_bench_steps(sim.h, args.steps, out)
    if rc != 0:
        raise RuntimeError(sim.lib.rd_last_error(sim.h).decode())
    wall_ms, gpu_ms, force_ms, sort_ms, _ = list(out)
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
        "kernel_avg_from": f"dispatch timestamps of the event pass (steps {args.warmup + args.steps}-"
                           f"{args.warmup + 2 * args.steps})",
        "bytes_alg_per_launch": B_ALG * n,
        "sort_avg_us": sort_ms / args.steps * 1e3,
    }
    passes = bench_passes(sim)
    sim.close()
    cfg_spec = dataclasses.replace(spec, cell_size=info.cell_size)
    full = configured_run(cfg_spec, xyz, types, lc)
    fperr = fp64_error()
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(spec, xyz, types)
    extras = {}
    if not args.no_extras:
        for name, fn in (("batched_replicas", bench_batched), ("learned_slab", bench_learned),
                         ("learned_membrane", bench_membrane), ("crowded_cytosol", bench_cytosol),
                         ("frame_handoff", bench_frames)):
            try:
                extras[name] = fn()
            except Exception as e:            # reported, not fatal: the headline line still prints
                log(f"{name} failed: {e}")
                extras[name] = {"error": str(e)}
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
            "force_kernel": {1: "gather", 2: "gather on a resident grid with work queues"}.get(lc["used_kernel"], "?"),
            "workgroup": lc["used_block"],
            "particles_per_lane": lc["used_ppt"],
            "timed_window": {"warmup": [0, args.warmup],
                             "autotune": "trial steps on the warmed state; state and step counter restored",
                             "graph_pass (value)": [args.warmup, args.warmup + args.steps],
                             "event_pass (roofline)": [args.warmup + args.steps, args.warmup + 2 * args.steps]},
        },
        "configured_run": full,
        "fp64_error": fperr,
        "roofline": roofline,
        "passes": passes,
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(res))


if __name__ == "__main__":
    main()

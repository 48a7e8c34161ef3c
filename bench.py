#!/usr/bin/env python3
"""Throughput of the device-resident DEM step on one MI355X.

Workload (default): the plain-method large settle -- 100,000 convex polyhedra
(hulls of 8-20 points on an ellipsoid 1.0/0.8/0.6 mm x U(0.8,1.2)) seeded on a
jittered lattice in a walled 200x200x400 mm box, falling under gravity; FP64;
dt = 1e-6 s; linear spring-dashpot / Coulomb law, Kn = 1e4, e = 0.5, mu = 0.3;
uniform cell hash with cell = 2 x max circumradius. Synthetic data (seeded
generator, no files).

A "step" is one full iteration of the reference time loop: wall slab test,
cell hash and candidate list, GJK/EPA and contact law on every candidate
pair, per-particle force/torque reduction and the Verlet integrator.
`value` is particle-steps/s with all inputs resident in HBM.

Prints one JSON line. --gpus N>1 is not supported (one GPU per box).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "particle-steps/s @100k polyhedra; batched vs plain; adjoint/forward cost"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def stage_bytes(sc_list, stats, nv_mean):
    """Algorithmic HBM bytes per launch of each stage."""
    M = sum(sc.n_particles for sc in sc_list)
    T = sum(sc.n_walls * sc.n_particles for sc in sc_list)
    P, K = stats["n_pairs"], stats["n_contacts"]
    V = nv_mean
    return {
        # walls: com read per test + flag; keys: com read + key/row write; sort: 3 passes x (8 in + 8 out)
        "walls+keys+sort": 24 * T + 4 * T + 24 * M + 8 * M + 3 * 16 * M,
        # cells + emission: sorted keys/rows read, pairs written
        "cells+emit": 8 * M + 8 * P,
        # GJK: pair ids, both centres, both vertex sets once, collision code
        "gjk": P * (8 + 48 + 2 * 24 * V + 4),
        "compact": 8 * P,
        # EPA + law per contact: ids 8, state of A and B 2*80, outputs 9*8*... (248 B/contact model)
        "epa+law": 248 * K,
        "record-sort": 2 * K * 16 * 2,
        # integrator: (84 + 6V) doubles of state per particle + records
        "reduce+verlet": M * (84 + 6 * V) * 8 + 2 * K * (4 + 48),
    }


def cpu_baseline(sc, budget_s: float = 12.0):
    """The oracle (FP64 C++ port of the reference path, one core) on the same
    scene, for as many whole steps as fit in about budget_s seconds."""
    from oracle import oracle as O
    st = O.initial_state(sc)
    t0 = time.perf_counter()
    n = 0
    while True:
        O.step(st, sc)
        n += 1
        if time.perf_counter() - t0 > budget_s or n >= 50:
            break
    dt = time.perf_counter() - t0
    return {"value": sc.n_particles * n / dt, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"{n} full steps of the same {sc.n_particles}-particle scene from its initial state "
                      f"(oracle/mwdiem_oracle.cpp, g++ -O2, one thread), {dt:.1f} s"}


def bench_batched(n_scenes: int, steps: int, plain_steps: int, graph_steps: int):
    """The batched form: n_scenes hopper-discharge scenes (particle counts
    log-uniform 500-20,000, per-scene friction, stiffness and restitution,
    physical law) in one Stepper -- one launch per stage over all scenes,
    graph-captured -- against the same scenes run one after another, each in
    its own graph-captured Stepper."""
    import torch

    from mwdiem_amd import scene
    from mwdiem_amd.stepper import Stepper

    t0 = time.perf_counter()
    scs = scene.config_batched(n_scenes=n_scenes)
    n_tot = sum(sc.n_particles for sc in scs)
    print(f"# batched: {n_scenes} scenes, {n_tot} particles, generated in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    B = Stepper(scs)
    B.step(5)
    B.check()
    B.capture(graph_steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    B.run(steps)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    sb = B.check()
    batched = n_tot * steps / tb
    B.close()
    del B
    torch.cuda.empty_cache()
    # plain: every scene alone, one after another
    t_plain = 0.0
    for sc in scs:
        S = Stepper(sc)
        S.step(2)
        S.capture(graph_steps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        S.run(plain_steps)
        torch.cuda.synchronize()
        t_plain += time.perf_counter() - t0
        S.check()
        S.close()
    plain = n_tot * plain_steps / t_plain
    print(f"# batched {batched:.4g} vs plain {plain:.4g} particle-steps/s", file=sys.stderr, flush=True)
    return {"scenes": n_scenes, "particles": n_tot, "steps": steps, "ms_per_step": 1e3 * tb / steps,
            "particle_steps_per_s": batched, "pairs_per_step": sb["n_pairs"], "contacts_per_step": sb["n_contacts"],
            "plain_steps_per_scene": plain_steps, "plain_particle_steps_per_s": plain, "batched_vs_plain": batched / plain}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--particles", type=int, default=100_000)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-graph", action="store_true", help="launch every step eagerly instead of graph replays")
    ap.add_argument("--graph-steps", type=int, default=20, help="steps captured per HIP graph")
    ap.add_argument("--batched-scenes", type=int, default=256)
    ap.add_argument("--batched-steps", type=int, default=1000)
    ap.add_argument("--plain-steps", type=int, default=100)
    ap.add_argument("--no-batched", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU design is "
              "described in DESIGN.md and is not measured.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from mwdiem_amd import scene
    from mwdiem_amd.stepper import Stepper

    t_gen = time.perf_counter()
    sc = scene.config_large(seed=args.seed, n=args.particles)
    t_gen = time.perf_counter() - t_gen
    print(f"# scene: {sc.n_particles} particles, vmax {sc.vmax}, cell {sc.max_particlesize:.3f} mm, "
          f"generated in {t_gen:.1f} s", file=sys.stderr, flush=True)
    S = Stepper(sc, device=0)
    for _ in range(args.warmup):
        S.step(1)
    S.check()
    torch.cuda.synchronize()
    if not args.no_graph:
        S.capture(args.graph_steps)   # records graph_steps steps; capture runs nothing


    # timed region: K full steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    S.run(args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    stats = S.check()
    # per-stage durations: HIP events between the stages on the stepper's
    # stream, over further eager steps (events cannot sit inside the graph)
    n_stage = min(args.steps, 500)
    S.set_timing(True)
    S.step(n_stage)
    st = S.stage_times()
    S.set_timing(False)
    ms_step = 1e3 * wall / args.steps
    value = sc.n_particles * args.steps / wall
    stage_ms = {k: v / max(st["steps"], 1) for k, v in st["ms"].items()}
    nv_mean = float(np.mean(sc.nverts[sc.n_walls:]))
    sbytes = stage_bytes([sc], stats, nv_mean)
    dom = max(stage_ms, key=stage_ms.get)
    ach = sbytes[dom] / (stage_ms[dom] * 1e-3) / 1e9
    roof = {"kernel": dom, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS, "traffic": None, "launch_ms": stage_ms[dom],
            "algorithmic_bytes": sbytes[dom]}
    print(f"# stages (ms/step): {json.dumps({k: round(v, 4) for k, v in stage_ms.items()})}", file=sys.stderr)
    print(f"# pairs {stats['n_pairs']} contacts {stats['n_contacts']} cells {stats['n_cells']}", file=sys.stderr,
          flush=True)

    S.close()
    del S
    torch.cuda.empty_cache()
    batched = None
    if not args.no_batched:
        batched = bench_batched(args.batched_scenes, args.batched_steps, args.plain_steps, args.graph_steps)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(sc, args.cpu_budget)

    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "plain large settle: 100,000 convex polyhedra (8-20 hull points), 200x200x400 mm "
                               "walled box, FP64, dt=1e-6 s, Kn=1e4, e=0.5, mu=0.3, cell=2*r_circ,max",
                   "particles": sc.n_particles, "pairs_per_step": stats["n_pairs"],
                   "contacts_per_step": stats["n_contacts"], "pair_tests_per_s": stats["n_pairs"] * args.steps / wall,
                   "stage_ms_per_step": {k: round(v, 5) for k, v in stage_ms.items()}},
        "roofline": roof,
        "cpu_baseline": cpu,
        "batched": batched,
        "batched_vs_plain": None if batched is None else batched["batched_vs_plain"],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

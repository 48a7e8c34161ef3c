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


# Algorithmic HBM bytes per launch of each single-kernel stage: what the
# kernel must read and write at least once (per-particle data reached
# through pair indices counted once per particle).
#  emit          : pair list out (8 B/pair), per cell its rank, start, 13
#                  neighbour ids and offset (4+8+52+8 B), members (4 B each)
#  sphere        : pair ids in (8 B), collision and flag out (8 B); centre
#                  and radius per row (32 B)
#  gjk           : per surviving pair its id (4), both rows (8), centres
#                  (48), vertex counts (8), both vertex sets (2 x 24 x V),
#                  collision/flag/slot out (12); a simplex (128) per contact
#  epa+law       : per contact: pair id, simplex, slot (4+128+4), both
#                  rows' centre, velocity, angular velocity, orientation and
#                  inertia size (2 x 112), vertex sets (2 x 24 x V), outputs
#                  depth/normal/points/force/torques (8+4x24+3x24) and two
#                  records (16)
#  reduce+verlet : per moving particle 448 B of state and constants read,
#                  272 B written, its vertices read and written (48 B each);
#                  per contact record 4 B index + 48 B force and torque; the
#                  record count of every row (4 B)
def stage_bytes(sc_list, stats):
    import numpy as np
    M = sum(sc.n_particles for sc in sc_list)
    N = sum(sc.n_rows for sc in sc_list)
    V = float(np.mean(np.concatenate([sc.nverts[sc.n_walls:] for sc in sc_list])))
    P, K, C = stats["n_pairs"], stats["n_contacts"], stats["n_cells"]
    S = stats["n_sphere_pass"]
    P_cell = P - stats["n_wall_pairs"]
    return {
        # emission with the bounding-sphere test folded in (pairs read back
        # from registers; 2 x 16 B centres and radii per pair, a collision
        # code and a survivor index written)
        "emit": 8 * P_cell + C * (4 + 8 + 52 + 8) + 4 * M + 16 * P + 32 * N + 4 * P + 4 * S,
        "sphere": 0,
        "gjk": S * (4 + 8 + 48 + 8 + 2 * 24 * V + 12) + 128 * K,
        "epa+law": K * (136 + 224 + 48 * V + 8 + 168 + 16),
        "reduce+verlet": M * (448 + 272 + 48 * V) + 2 * K * 52 + 4 * N,
    }


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` from the committed PMC summary for this
    workload (profiles/<kernel>_<workload>_pmc.md, line 'hbm_bytes_per_launch: N'),
    or None when there is none."""
    f = ROOT / "profiles" / f"{kernel}_{workload}_pmc.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("hbm_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def host_cpu() -> dict:
    """Host cores as the CPU baseline sees them: logical CPUs of the machine,
    those this process may run on, and the model name."""
    model = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "model": model}


def cpu_baseline(sc, budget_s: float = 12.0, gjk_samples: dict | None = None):
    """The oracle (the FP64 C++ restatement of the reference path, OpenMP over
    pairs and particles, contact sums in pair order) on the same scene, for
    as many whole steps as fit in about budget_s seconds. It also counts, with
    its instrumented double (add/sub/mul/div/sqrt = 1 flop), the FP64 work of
    GJK, EPA and the contact law on each of gjk_samples' pair lists -- the
    pairs the GPU's GJK launch processed (sphere-test survivors) and their
    state -- the algorithmic flops of that launch."""
    import ctypes

    from oracle import oracle as O
    O.lib().orc_threads.restype = ctypes.c_int
    threads = int(O.lib().orc_threads())
    st = O.initial_state(sc)
    t0 = time.perf_counter()
    n = 0
    while True:
        O.step(st, sc)
        n += 1
        if time.perf_counter() - t0 > budget_s or n >= 200:
            break
    dt = time.perf_counter() - t0
    out = {"value": sc.n_particles * n / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
           "sample": f"{n} full steps of the same {sc.n_particles}-particle scene from its initial state "
                     f"(oracle/mwdiem_oracle.cpp, g++ -O2 -fopenmp, {threads} threads), {dt:.1f} s",
           "host": host_cpu()}
    flops = {}
    for name, smp in (gjk_samples or {}).items():
        t1 = time.perf_counter()
        flops[name] = O.count_flops(smp["pairs"], smp["state"], smp["scene"])
        print(f"# flops of {name}: {flops[name]} over {len(smp['pairs'])} pairs (the GPU's survivors: "
              f"{smp.get('survivors_on_gpu')}; {time.perf_counter() - t1:.1f} s)", file=sys.stderr, flush=True)
    return out, flops


def cpu_baseline_python(sc, budget_s: float = 10.0):
    """The reference's host loop in pure Python (oracle/ref_structure.py: dict
    and set cell hash, per-pair gather, each pair's GJK/EPA and law one after
    another, each particle's Verlet step), one core, on the small scene from
    its initial state, for as many whole steps as fit in about budget_s."""
    from oracle import oracle as O
    from oracle import ref_structure as R
    st = O.initial_state(sc)
    t0 = time.perf_counter()
    n = 0
    while True:
        R.step(st, sc)
        n += 1
        if time.perf_counter() - t0 > budget_s or n >= 1000:
            break
    dt = time.perf_counter() - t0
    return {"value": sc.n_particles * n / dt, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"{n} of the 1,000 steps of the small scene ({sc.n_particles} hulls of 12 points), the "
                      f"reference's host loop restated in pure Python (oracle/ref_structure.py), one core, "
                      f"{dt:.1f} s", "host": host_cpu()}


def gjk_sample(S, sc):
    """The pair list the next step's GJK launch takes (candidate pairs of the
    current state whose bounding spheres meet; wall pairs whose particle's
    bounding sphere and wall slab are not apart by more than 1e-6 (1 + |d| + h),
    the wall scan's test), with the state, for counting that launch's
    algorithmic flops. Its length equals the step's sphere-test survivor count."""
    import numpy as np
    S.candidates()
    stt = S.check()
    pairs = S.pairs()
    st = S.numpy_state()
    r = S.t["radius"].cpu().numpy()
    nw = stt["n_wall_pairs"]
    cell = pairs[nw:]
    d = st["com"][cell[:, 1]] - st["com"][cell[:, 0]]
    meet = np.einsum("ij,ij->i", d, d) <= (r[cell[:, 0]] + r[cell[:, 1]]) ** 2
    wall = pairs[:nw]
    j, k = wall[:, 0], wall[:, 1]
    wc = S.t["wall_c"].cpu().numpy() if "wall_c" in S.t else sc.wall_c
    n = sc.wall_n[k]
    dd = np.abs(np.einsum("ij,ij->i", n, st["com"][j] - wc[k]))
    h = np.zeros(nw)
    for w in np.unique(k):
        V = st["verts"][w, :sc.nverts[w]] - wc[w]
        h[k == w] = np.abs(V @ sc.wall_n[w]).max()
    keep = ~(dd - h - r[j] > 1e-6 * (1.0 + dd + h))
    surv = np.concatenate([wall[keep], cell[meet]]).astype(np.int32)
    return {"pairs": surv, "state": st, "scene": sc, "candidates": int(pairs.shape[0]),
            "survivors_on_gpu": int(stt["n_sphere_pass"])}


def roofline_fp64(flops, launch_ms, kernel, workload, timed_over):
    """FP64 roofline of the GJK launch: algorithmic flops per launch over its
    mean duration against MI355X's FP64 vector peak; the PMC figures of the
    same launch (FP64 VALU instructions per launch) when a summary exists."""
    if not flops or not launch_ms:
        return None
    ach = flops / (launch_ms * 1e-3) / 1e12
    return {"bound": "fp64_valu", "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": ach / FP64_PEAK_TFLOPS, "algorithmic_flops": flops, "kernel": kernel, "launch_ms": launch_ms,
            "timed_over": timed_over, "pmc": measured_fp64(kernel, workload)}


def measured_fp64(kernel: str, workload: str):
    """FP64 VALU instructions per launch of `kernel` from the committed PMC
    summary profiles/<kernel>_<workload>_fp64_pmc.md (lines 'name: value')."""
    f = ROOT / "profiles" / f"{kernel}_{workload}_fp64_pmc.md"
    if not f.exists():
        return None
    out = {}
    for line in f.read_text().splitlines():
        for key in ("fp64_instructions_per_launch", "fp64_lane_flops_per_launch", "dispatches"):
            if line.startswith(key + ":"):
                out[key] = float(line.split(":", 1)[1])
    return out or None


def plain_leg(S, sc, args, workload: str, label: str, tune_space=None):
    """One plain-method measurement on a Stepper: warm-up, launch choices by
    the autotuner, kernel durations over eager steps with start/stop events,
    the pair list of the next GJK launch (for its flop count), then K
    graph-replayed steps timed."""
    import torch

    from mwdiem_amd import _lib
    for _ in range(args.warmup):
        S.step(1)
    S.check()
    tuned = None
    if not args.no_tune:
        # launch choices by timing (results are bit-identical across them;
        # the state is restored after the trials)
        from mwdiem_amd.tune import autotune
        tuned = autotune(S, steps=10, space=tune_space)
        print(f"# {label} autotune: {tuned['tuning']} {tuned['ms_per_step']:.4f} ms/step (defaults "
              f"{tuned['default_ms_per_step']:.4f}), {len(tuned['trials'])} trials", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    # Kernel durations: eager steps on the stepper's stream with HIP events --
    # stage boundaries, and for the single-kernel stages start/stop events
    # stamped with the kernel's own start and end (hipExtLaunchKernelGGL), so
    # launch gaps are excluded.
    n_time = min(args.steps, args.timing_steps)
    S.set_timing(True)
    S.step(n_time)
    kt = S.kernel_times()
    st = S.stage_times()
    stats_t = S.check()  # counts of the last timed-kernel step
    S.set_timing(False)
    timing_src = f"{kt['steps']} eager steps after warm-up, kernel start/stop events"
    sample = gjk_sample(S, sc)
    if not args.no_graph:
        S.capture(args.graph_steps)   # records graph_steps steps; capture runs nothing
        S.run(args.graph_steps)       # an untimed replay: a graph's first launch also uploads it
    # timed region: K full steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    S.run(args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    stats = S.check()
    stage_ms = {k: v / max(st["steps"], 1) for k, v in st["ms"].items()}
    kernel_ms = {k: v / max(kt["steps"], 1) for k, v in kt["ms"].items()}
    sbytes = stage_bytes([sc], stats_t)
    stage_of = dict(zip(_lib.KTIMED_NAMES, _lib.KTIMED_STAGE))
    # the dominant kernel: the slowest of the single-kernel stages
    kname = max(kernel_ms, key=kernel_ms.get)
    dom = stage_of[kname]
    ach = sbytes[dom] / (kernel_ms[kname] * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": measured_traffic(kname, workload), "kernel": kname, "launch_ms": kernel_ms[kname],
            "algorithmic_bytes": sbytes[dom], "timed_over": timing_src,
            "kernel_ms": {k: round(v, 5) for k, v in kernel_ms.items()},
            "kernel_gbs": {k: round(sbytes[stage_of[k]] / (v * 1e-3) / 1e9, 1) for k, v in kernel_ms.items() if v > 0}}
    print(f"# {label} stages (ms/step): {json.dumps({k: round(v, 4) for k, v in stage_ms.items()})}",
          file=sys.stderr)
    print(f"# {label}: pairs {stats['n_pairs']} GJK pairs {stats['n_sphere_pass']} contacts {stats['n_contacts']} "
          f"cells {stats['n_cells']}; {1e3 * wall / args.steps:.4f} ms/step", file=sys.stderr, flush=True)
    return {"wall": wall, "stats": stats, "stats_timed": stats_t, "stage_ms": stage_ms, "kernel_ms": kernel_ms,
            "roofline": roof, "tuned": tuned, "timing_src": timing_src, "sample": sample}


def profile_leg(args):
    """One workload stepped eagerly, for a rocprofv3 run: --warmup steps,
    then --steps steps, default launch choices, then the counts."""
    import numpy as np

    from mwdiem_amd import scene
    from mwdiem_amd.stepper import Stepper
    if args.profile_leg == "adjoint":
        profile_adjoint(args)
        return
    kw = {}
    v = None
    if args.profile_leg == "poly":
        sc = scene.config_polydisperse()
        rng = np.random.Generator(np.random.PCG64(3))
        v = np.zeros((sc.n_rows, 3))
        v[sc.n_walls:] = rng.normal(size=(sc.n_particles, 3)) * 1000.0
        v[:sc.n_walls] = sc.wall_v[:sc.n_walls]
    else:
        sc = scene.config_large(seed=args.seed, n=args.particles)
        if args.profile_leg == "contacting":
            sc = contacting_scene(sc)
            kw["pair_cap"] = 20_000_000
    S = Stepper(sc, **kw)
    if v is not None:
        S.set_state({"v": v, "v12": v})
    print(f"# profile {args.profile_leg}: {sc.n_particles} hulls, {args.warmup} + {args.steps} steps",
          file=sys.stderr, flush=True)
    S.step(args.warmup)
    S.check()
    S.step(args.steps)
    print(json.dumps({"profile_leg": args.profile_leg, "steps": args.steps, "warmup": args.warmup,
                      "status": S.check()}))
    S.close()


def profile_adjoint(args):
    """The differentiable form's reverse sweep for a rocprofv3 run: the
    4,096-hull pile poured on its spawn schedule, then the bed outcome's
    gradient over --steps steps by the adjoint with checkpoints every 100
    steps and the tape, graphs of --graph-steps steps, fixed launch choices
    for a packed bed (the sorted cell pass, 4 cells per fill block shared by
    16 blocks, 4 lanes per GJK pair)."""
    import numpy as np

    from mwdiem_amd import scene
    from mwdiem_amd.adjoint import adjoint_run
    from mwdiem_amd.diff import form_pile
    sc, v_pour, spawn = scene.config_pour(n=4096)
    n_pour = spawn.every * ((sc.n_particles + spawn.per - 1) // spawn.per) + 40_000
    tune = dict(fill_cells=4, gjk_group=4)
    print(f"# profile adjoint: pouring {sc.n_particles} hulls over {n_pour} steps", file=sys.stderr, flush=True)
    st0 = form_pile(sc, v_pour, steps=n_pour, spawn=spawn, steps_per_graph=args.graph_steps, tuning=tune)
    tune.update(cell_pass=1, fill_split=16)  # the sorted pass (the pour's spawn schedule needs the key tables)
    theta = np.array([sc.law_p.Kn, sc.law_p.en, sc.law_p.mu])
    print(f"# profile adjoint: {args.steps} steps forward, then the reverse sweep", file=sys.stderr, flush=True)
    tm = {}
    L, g = adjoint_run(sc, st0, theta, args.steps, every=100, steps_per_graph=args.graph_steps, timings=tm,
                       tuning=tune)
    print(json.dumps({"profile_leg": "adjoint", "steps": args.steps, "outcome": L, "grad": [float(x) for x in g],
                      "forward_ms_per_step": 1e3 * tm["forward_s"] / args.steps,
                      "reverse_ms_per_step": 1e3 * tm["reverse_s"] / args.steps,
                      "outcome_and_seed_ms": 1e3 * tm["outcome_s"]}))


def contacting_scene(sc, gap: float = -0.6):
    """A contact-bearing state of the large settle's hulls: the same 100,000
    hulls moved onto a body-centred-cubic lattice from the floor up whose
    spacing is (1 + gap) times twice the largest circumradius, so that
    neighbours interpenetrate (about 1.5 contacts per hull, 13 M candidate
    pairs), on the physical law (the reference law's damping is unstable once
    hulls touch)."""
    from mwdiem_amd import scene
    out = scene.compact_dense(sc, gap=gap)
    out.law_p = scene.Law4(Kn=sc.law_p.Kn, en=sc.law_p.en, et=sc.law_p.et, mu=sc.law_p.mu, mode=scene.PHYSICAL)
    out.law_w = scene.Law4(Kn=sc.law_w.Kn, en=sc.law_w.en, et=sc.law_w.et, mu=sc.law_w.mu, mode=scene.PHYSICAL)
    out.name = f"{sc.name}-compacted"
    return out


def bench_batched(scs, steps: int, plain_steps: int, graph_steps: int, tune: bool = True, v_rms: float = 5000.0):
    """The batched form: n_scenes hopper-discharge scenes (particle counts
    log-uniform 500-20,000, per-scene friction, stiffness and restitution,
    physical law) in one Stepper -- one launch per stage over all scenes,
    graph-captured, cell size and launch choices picked by the autotuner --
    against the same scenes (same cells) run one after another, each in its
    own graph-captured Stepper with the default launch choices. Every scene
    starts with its hulls thrown about (v_rms mm/s per component, seeded per
    scene), so contacts form from the first steps in every scene and GJK/EPA
    work is uneven across them; the launch choices are timed 200 steps into
    the throw, and kernel durations of the batched step come from the
    library's kernel timers over 20 eager steps after the timed run."""
    import numpy as np
    import torch

    from mwdiem_amd.stepper import Stepper
    from mwdiem_amd.tune import autotune, regrid, tune_cell_scale

    n_scenes = len(scs)
    n_tot = sum(sc.n_particles for sc in scs)
    print(f"# batched: {n_scenes} scenes, {n_tot} particles", file=sys.stderr, flush=True)
    vel = []
    for i, sc in enumerate(scs):
        v = np.zeros((sc.n_rows, 3))
        v[sc.n_walls:] = np.random.Generator(np.random.PCG64(1000 + i)).normal(size=(sc.n_particles, 3)) * v_rms
        vel.append(v)
    v_all = np.concatenate(vel)
    tuned = None
    if tune:
        scale, scale_ms = tune_cell_scale(scs, steps=5)
        scs = [regrid(sc, scale) for sc in scs]
        print(f"# batched cell scale {scale} ({scale_ms})", file=sys.stderr, flush=True)
    B = Stepper(scs)
    B.set_state({"v": v_all, "v12": v_all})
    B.step(5)
    B.check()
    if tune:
        # launch choices timed on the state 200 steps on, once the throw has
        # made contacts in every scene (GJK's choices matter only then); the
        # timed run starts from step 5, as the plain runs do
        snap = B.snapshot()
        B.step(200)
        tuned = autotune(B, steps=10)
        B.restore(snap)
        tuned = {"cell_scale": scale, "cell_scale_ms_per_step": scale_ms, "tuning": tuned["tuning"],
                 "ms_per_step_eager": tuned["ms_per_step"], "default_ms_per_step_eager": tuned["default_ms_per_step"]}
        print(f"# batched autotune: {tuned['tuning']}", file=sys.stderr, flush=True)
    B.capture(graph_steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    B.run(steps)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    sb = B.check()
    batched = n_tot * steps / tb
    B.set_timing(True)
    B.step(20)
    kt = B.kernel_times()
    B.stage_times()
    B.set_timing(False)
    kernel_ms = {k: v / max(kt["steps"], 1) for k, v in kt["ms"].items()}
    sb_end = B.check()
    print(f"# batched kernels (ms/step, library timers, after {steps} steps, {sb_end['n_contacts']} contacts): "
          f"{kernel_ms}", file=sys.stderr, flush=True)
    B.close()
    del B
    torch.cuda.empty_cache()
    # plain: every scene alone, one after another, from the same throw
    t_plain = 0.0
    for sc, v in zip(scs, vel):
        S = Stepper(sc)
        S.set_state({"v": v, "v12": v})
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
            "v_rms_mm_s": v_rms, "kernel_ms_at_end": kernel_ms,
            "plain_steps_per_scene": plain_steps, "plain_particle_steps_per_s": plain, "batched_vs_plain": batched / plain,
            "autotune": tuned}


def fd_richardson(outcome, theta, k, h_rel=1e-4):
    """Central differences of outcome(theta) in parameter k at h, h/2 and h/4
    (h = h_rel * theta_k) and their Richardson extrapolations (the central
    difference's error is c h^2 + O(h^4)): returns (D(h), D(h/2), D(h/4),
    R(h, h/2), R(h/2, h/4))."""
    import numpy as np
    D = []
    for f in (1.0, 0.5, 0.25):
        h = h_rel * f * theta[k]
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        D.append((outcome(tp) - outcome(tm)) / (2 * h))
    return D + [(4 * D[1] - D[0]) / 3, (4 * D[2] - D[1]) / 3]


def bench_diff(pour, steps: int, graph_steps: int, horizons=(10, 100, 1000, 2000)):
    """The differentiable form on the poured pile: the pour configuration's
    4,096 hulls enter a 40x40x60 mm box on their spawn schedule and settle
    (physical law Kn=1e4, en=0.5, mu=0.3); from that pile, the cost of a
    tangent step (state plus its derivatives with respect to Kn, en, mu) and
    of the adjoint reverse sweep (checkpoints every 100 steps) against a plain
    step over `steps` steps, graph-captured; and the bed outcome's gradient
    (tangents) against central differences at h = 1e-4, 5e-5 and 2.5e-5
    relative with Richardson extrapolation, at several horizons. The longest
    horizon at which the differences converge (to 1e-6) and agree with the
    gradient (to 1e-5) is reported; on the settled pile there is none (its
    outcome moves by jumps as contacts among ~8,000 open and close), and the
    gradient's check is the PyTorch-CPU reference in the GPU tests."""
    import numpy as np
    import torch

    from mwdiem_amd.diff import _with_theta, form_pile, run_outcome
    from mwdiem_amd.stepper import Stepper

    from mwdiem_amd.tune import autotune

    sc, v_pour, spawn = pour
    n = sc.n_particles
    W = sc.n_walls
    n_pour = spawn.every * ((n + spawn.per - 1) // spawn.per) + 40_000
    t0 = time.perf_counter()
    # launch choices for a packed bed of few cells (results do not depend on them)
    st0 = form_pile(sc, v_pour, steps=n_pour, spawn=spawn, steps_per_graph=graph_steps,
                    tuning=dict(fill_cells=4, gjk_group=4))
    t_pour = time.perf_counter() - t0
    speed = np.linalg.norm(st0["v"][W:], axis=1)
    pile = {"pour_steps": n_pour, "pour_s": t_pour, "batches": (n + spawn.per - 1) // spawn.per,
            "per_batch": spawn.per, "batch_every_steps": spawn.every,
            "median_speed_mm_s": float(np.median(speed)), "max_speed_mm_s": float(speed.max()),
            "kinetic_energy": float(0.5 * np.sum(sc.mass[W:] * speed ** 2)),
            "top10_mean_z_mm": float(np.sort(st0["com"][W:, 2])[-n // 10:].mean()),
            "finite": bool(all(np.isfinite(a).all() for a in st0.values()))}
    print(f"# diff: pile poured in {n_pour} steps ({t_pour:.1f} s): {pile}", file=sys.stderr, flush=True)
    theta = np.array([sc.law_p.Kn, sc.law_p.en, sc.law_p.mu])
    S = Stepper(_with_theta(sc, theta))
    S.set_state(st0)
    tune = autotune(S, steps=10)["tuning"]
    S.close()
    print(f"# diff: pile autotune {tune}", file=sys.stderr, flush=True)
    times = {}
    for mode in ("plain", "tangent"):
        S = Stepper(_with_theta(sc, theta))
        S.set_tuning(**tune)
        S.set_state(st0)
        if mode == "tangent":
            S.enable_tangents()
        S.capture(graph_steps, tangent=(mode == "tangent"))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        S.run(steps)
        torch.cuda.synchronize()
        times[mode] = (time.perf_counter() - t0) / steps
        st = S.check()
        S.close()
    print(f"# diff: tangent {1e3 * times['tangent']:.4f} ms/step vs plain {1e3 * times['plain']:.4f}; "
          f"{st['n_contacts']} contacts", file=sys.stderr, flush=True)
    # reverse mode: the gradient by the adjoint, checkpoints every 100 steps
    # (forward, then per segment a recomputation and the reverse sweep)
    from mwdiem_amd.adjoint import adjoint_run
    from mwdiem_amd.diff import bed_outcome
    every = 100
    tm = {}
    # torch loads the kernels of the outcome (top-k, reductions) and of its
    # autograd at their first call in the process (about 0.5 s): one untimed
    # call on a tensor of the same shape first
    with torch.enable_grad():
        x = torch.zeros((sc.n_rows, 3), dtype=torch.float64, device="cuda").requires_grad_(True)
        torch.autograd.grad(bed_outcome(x, W), x)
    torch.cuda.synchronize()
    La, ga = adjoint_run(sc, st0, theta, steps, every=every, steps_per_graph=graph_steps, timings=tm, tuning=tune)
    checks = {}
    longest = None
    for H in horizons:
        L, g = run_outcome(sc, st0, theta, H, True, graph_steps, tuning=tune)
        out = {"outcome": L, "grad": [float(x) for x in g]}

        def outcome(th, H=H):
            return run_outcome(sc, st0, th, H, False, graph_steps, tuning=tune)[0]
        conv, rel, fd = [], [], []
        for k in range(3):
            D0, D1, D2, R1, R2 = fd_richardson(outcome, theta, k)
            fd.append([D0, D1, D2, R2])
            conv.append(abs(R2 - R1) / max(abs(R2), 1e-300))
            rel.append(abs(g[k] - R2) / max(abs(R2), 1e-300))
        out.update({"fd_h_h2_h4_richardson": fd, "fd_convergence": conv, "rel_to_richardson": rel})
        ok = max(conv) <= 1e-6 and max(rel) <= 1e-5
        out["agrees"] = bool(ok)
        if ok:
            longest = H
        checks[str(H)] = out
        print(f"# diff: horizon {H}: conv {['%.1e' % c for c in conv]} rel {['%.1e' % r for r in rel]}",
              file=sys.stderr, flush=True)
    if steps == horizons[-1]:
        gt = np.array(checks[str(steps)]["grad"])
    else:
        gt = run_outcome(sc, st0, theta, steps, True, graph_steps, tuning=tune)[1]
    adj = {"checkpoint_every": every, "outcome": La, "grad": [float(x) for x in ga],
           "rel_to_tangents": [float(x) for x in np.abs(ga - gt) / np.maximum(np.abs(gt), 1e-300)],
           "forward_ms_per_step": 1e3 * tm["forward_s"] / steps,
           "reverse_ms_per_step": 1e3 * tm["reverse_s"] / steps,
           "outcome_and_seed_ms": 1e3 * tm["outcome_s"],
           "adjoint_vs_forward": tm["reverse_s"] / steps / times["plain"],
           "gradient_vs_forward": (tm["forward_s"] + tm["outcome_s"] + tm["reverse_s"]) / steps / times["plain"]}
    print(f"# diff: adjoint {adj}", file=sys.stderr, flush=True)
    return {"particles": n, "steps": steps, "pile": pile, "ms_per_step_plain": 1e3 * times["plain"],
            "ms_per_step_tangent": 1e3 * times["tangent"], "tangent_vs_forward": times["tangent"] / times["plain"],
            "adjoint": adj, "adjoint_vs_forward": adj["adjoint_vs_forward"], "gradient_checks": checks,
            "longest_horizon_fd_agrees": longest, "contacts_last_step": st["n_contacts"],
            "theta": ["Kn", "en", "mu"],
            "mode": "forward tangents (3 directions); reverse adjoint with step checkpoints"}


def bench_polydisperse(sc, steps: int, graph_steps: int, v_rms: float = 1000.0, tune: bool = True):
    """The polydisperse stress configuration: 1,000,000 hulls (circumradius
    log-uniform 0.5-2.5 mm, 6-32 points) in a 400 mm box whose ceiling slides
    along x at 0.1 m/s, physical law, FP64 GJK/EPA. The hulls start from
    their lattice with random velocities (v_rms mm/s per component, seeded),
    so that they collide unevenly; `steps` graph-replayed steps with the FP64
    and with the FP32 contact reduction, and the largest position difference
    between the two runs."""
    import numpy as np
    import torch

    from mwdiem_amd.stepper import Stepper

    print(f"# polydisperse: {sc.n_particles} hulls, vmax {sc.vmax}", file=sys.stderr, flush=True)
    W = sc.n_walls
    rng = np.random.Generator(np.random.PCG64(3))
    v = np.zeros((sc.n_rows, 3))
    v[W:] = rng.normal(size=(sc.n_particles, 3)) * v_rms
    v[:W] = sc.wall_v[:W]
    out, com, speed = {}, {}, {}
    tuning = None
    if tune:
        # launch choices timed half-way through the run, once the bed has
        # lost its lattice order and the contacts have formed (on the first
        # steps there are none, and the cell-pass and EPA choices timed there
        # do not hold later), then used for both timed runs from the start
        from mwdiem_amd.tune import SPACE, autotune
        S = Stepper(sc)
        S.set_state({"v": v, "v12": v})
        S.step(max(2, steps // 2))
        tuning = autotune(S, steps=5, space=dict(SPACE, epa_fused=(0, 1)))["tuning"]
        print(f"# polydisperse autotune (after {max(2, steps // 2)} steps, {S.status()['n_contacts']} contacts): "
              f"{tuning}", file=sys.stderr, flush=True)
        S.close()
        del S
        torch.cuda.empty_cache()
    for fp32 in (False, True):
        S = Stepper(sc, reduce_fp32=fp32)
        S.set_state({"v": v, "v12": v})
        S.step(2)
        if tuning is not None:
            S.set_tuning(**tuning)
        S.capture(graph_steps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        S.run(steps)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        st = S.check()
        fin = S.numpy_state()
        if not fp32:
            # kernel durations at the end of the run (contacts formed) and the
            # pair list of the next GJK launch, for its flop count
            S.set_timing(True)
            S.step(20)
            kt = S.kernel_times()
            S.stage_times()
            S.set_timing(False)
            kernel_ms = {k: v / max(kt["steps"], 1) for k, v in kt["ms"].items()}
            sample = gjk_sample(S, sc)
            sample["kernel_ms"] = kernel_ms
            sample["stats"] = S.check()
        com[fp32] = fin["com"][W:]
        speed[fp32] = np.linalg.norm(fin["v"][W:], axis=1)
        ke = 0.5 * float(np.sum(sc.mass[W:] * speed[fp32] ** 2))
        out["fp32" if fp32 else "fp64"] = {"ms_per_step": 1e3 * dt / steps,
                                           "particle_steps_per_s": sc.n_particles * steps / dt,
                                           "pairs_per_step": st["n_pairs"], "contacts_per_step": st["n_contacts"],
                                           "kinetic_energy": ke,
                                           "runaway_hulls": int(np.sum(speed[fp32] > 10 * v_rms))}
        S.close()
        del S
        torch.cuda.empty_cache()
    disp = float(np.abs(com[False] - sc.com[sc.n_walls:]).max())
    # a granular gas is chaotic: compare the two reductions statistically
    # (kinetic energy, contact count, how many hulls have parted company)
    # (hulls faster than 10 v_rms in either run -- the smallest, stiffest ones
    # at this dt, see the stability note -- are left out of the energy check)
    dx = np.linalg.norm(com[True] - com[False], axis=1)
    calm = (speed[True] <= 10 * v_rms) & (speed[False] <= 10 * v_rms)
    m = sc.mass[W:]
    ke_calm = [0.5 * float(np.sum(m[calm] * speed[f][calm] ** 2)) for f in (False, True)]
    out.update({"particles": sc.n_particles, "steps": steps + 2, "v_rms_mm_s": v_rms, "tuning": tuning,
                "kinetic_energy_rel_diff": abs(out["fp32"]["kinetic_energy"] / out["fp64"]["kinetic_energy"] - 1.0),
                "calm_hulls": int(calm.sum()), "calm_kinetic_energy_rel_diff": abs(ke_calm[1] / ke_calm[0] - 1.0),
                "median_dx_fp32_vs_fp64_mm": float(np.median(dx)), "frac_dx_above_1um": float(np.mean(dx > 1e-3)),
                "max_dx_fp32_vs_fp64_mm": float(dx.max()), "max_displacement_mm": disp})
    out["kernel_ms_at_end"] = {k: round(v, 5) for k, v in sample["kernel_ms"].items()}
    out["gjk_pairs_at_end"] = int(sample["pairs"].shape[0])
    print(f"# polydisperse: {out}", file=sys.stderr, flush=True)
    return out, sample


def copy_probe(nbytes: int, reps: int = 5):
    """A plain device copy of nbytes, reps times: the calibration run for the
    FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "reps": reps}))


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
    ap.add_argument("--batched-steps", type=int, default=5000)
    ap.add_argument("--plain-steps", type=int, default=100)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--timing-steps", type=int, default=200, help="eager steps timed per kernel with events")
    ap.add_argument("--diff-steps", type=int, default=2000, help="steps of the differentiable-form measurement")
    ap.add_argument("--no-diff", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch choices (no autotuner)")
    ap.add_argument("--poly-steps", type=int, default=2000, help="steps of the polydisperse stress configuration")
    ap.add_argument("--no-poly", action="store_true")
    ap.add_argument("--no-contacting", action="store_true", help="skip the compacted (contact-bearing) C2 bed")
    ap.add_argument("--profile-leg", choices=("c2", "contacting", "poly", "adjoint"), default=None,
                    help="only step one workload eagerly (warm-up plus --steps steps) for a profiler run: the "
                         "large settle, its compacted bed, or the polydisperse configuration thrown about at "
                         "1 m/s (--warmup steps first, then --steps steps); or 'adjoint': the poured pile's "
                         "reverse sweep over --steps steps")
    ap.add_argument("--copy-probe", type=int, default=0, help="only run a device copy of this many bytes (counter "
                                                                 "calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): this build measures a single "
              "MI355X; sharding across GPUs is a design, not a measured path.", file=sys.stderr)
        sys.exit(2)

    # every scene is made before the GPU runtime starts in this process
    # (the generator forks worker processes only before that)
    from mwdiem_amd import scene
    if args.profile_leg:
        profile_leg(args)
        return
    t_gen = time.perf_counter()
    sc = scene.config_large(seed=args.seed, n=args.particles)
    print(f"# scene: {sc.n_particles} particles, vmax {sc.vmax}, cell {sc.max_particlesize:.3f} mm, "
          f"generated in {time.perf_counter() - t_gen:.1f} s", file=sys.stderr, flush=True)
    sc_dense = None if args.no_contacting else contacting_scene(sc)
    scs_batched = None if args.no_batched else scene.config_batched(n_scenes=args.batched_scenes)
    pour = None if args.no_diff else scene.config_pour(n=4096)
    sc_poly = None if args.no_poly else scene.config_polydisperse()
    sc_small = None if args.no_cpu_baseline else scene.config_small()
    print(f"# all scenes generated in {time.perf_counter() - t_gen:.1f} s", file=sys.stderr, flush=True)

    import torch

    from mwdiem_amd.stepper import Stepper
    from mwdiem_amd.tune import SPACE

    S = Stepper(sc, device=0)
    main_leg = plain_leg(S, sc, args, "c2_100k", "C2")
    S.close()
    del S
    torch.cuda.empty_cache()
    wall, stats = main_leg["wall"], main_leg["stats"]
    value = sc.n_particles * args.steps / wall
    samples = {"c2_100k": main_leg["sample"]}

    contacting = None
    if sc_dense is not None:
        S = Stepper(sc_dense, device=0, pair_cap=20_000_000)
        leg = plain_leg(S, sc_dense, args, "c2_100k_contacting", "C2 compacted",
                        tune_space=dict(SPACE, epa_fused=(0, 1)))
        S.close()
        del S
        torch.cuda.empty_cache()
        samples["c2_100k_contacting"] = leg["sample"]
        contacting = {"workload": "the large settle's 100,000 hulls compacted onto a lattice 0.4x twice the largest "
                                  "circumradius (neighbours interpenetrate), physical law, same Kn, e, mu, dt",
                      "ms_per_step": 1e3 * leg["wall"] / args.steps,
                      "particle_steps_per_s": sc_dense.n_particles * args.steps / leg["wall"],
                      "candidates_per_step": leg["stats"]["n_pairs"],
                      "gjk_pairs_per_step": leg["stats"]["n_sphere_pass"],
                      "contacts_per_step": leg["stats"]["n_contacts"],
                      "pair_tests_per_s": leg["stats"]["n_sphere_pass"] * args.steps / leg["wall"],
                      "stage_ms_per_step": {k: round(v, 5) for k, v in leg["stage_ms"].items()},
                      "roofline": leg["roofline"], "tuning": None if leg["tuned"] is None else leg["tuned"]["tuning"],
                      "kernel_ms": leg["kernel_ms"]}

    batched = None
    if scs_batched is not None:
        batched = bench_batched(scs_batched, args.batched_steps, args.plain_steps, args.graph_steps,
                                tune=not args.no_tune)

    diff = None
    if pour is not None:
        diff = bench_diff(pour, args.diff_steps, args.graph_steps)

    poly = None
    if sc_poly is not None:
        poly, samples["c3_1m"] = bench_polydisperse(sc_poly, args.poly_steps, args.graph_steps,
                                                    tune=not args.no_tune)

    cpu, flops, cpu_py = None, {}, None
    if not args.no_cpu_baseline:
        cpu, flops = cpu_baseline(sc, args.cpu_budget, samples)
        cpu_py = cpu_baseline_python(sc_small, args.cpu_budget)
        print(f"# cpu baselines: {cpu['value']:.4g} (oracle, {cpu['cores']} threads), {cpu_py['value']:.4g} "
              f"(reference structure in Python, 1 core) particle-steps/s", file=sys.stderr, flush=True)

    # FP64 roofline of the GJK launch (GJK, and with epa_fused EPA and the law)
    from mwdiem_amd import _lib
    gjk_name = _lib.KTIMED_NAMES[_lib.KTIMED_STAGE.index("gjk")]
    ms = main_leg["kernel_ms"]
    roof_gjk = {"c2_100k": roofline_fp64(flops.get("c2_100k"), ms.get(gjk_name, 0.0) + ms.get("k_epa_force", 0.0),
                                         "k_gjk_epa", "c2_100k", main_leg["timing_src"])}
    if contacting is not None:
        ms = contacting["kernel_ms"]
        # with EPA in its own launch the GJK+EPA work spans both kernels
        both = ms.get(gjk_name, 0.0) + ms.get("k_epa_force", 0.0)
        roof_gjk["c2_100k_contacting"] = roofline_fp64(flops.get("c2_100k_contacting"), both, "k_gjk_epa",
                                                       "c2_100k_contacting", "eager steps, kernel start/stop events")
        contacting["kernel_ms"] = {k: round(v, 5) for k, v in ms.items()}
    if poly is not None:
        ms = samples["c3_1m"]["kernel_ms"]
        both = ms.get(gjk_name, 0.0) + ms.get("k_epa_force", 0.0)
        roof_gjk["c3_1m"] = roofline_fp64(flops.get("c3_1m"), both, "k_gjk_epa", "c3_1m",
                                          "20 eager steps after the timed run, kernel start/stop events")
        poly["gjk_pairs_per_s_at_end"] = samples["c3_1m"]["stats"]["n_sphere_pass"] / (both * 1e-3) if both else None

    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "plain large settle: 100,000 convex polyhedra (8-20 hull points), 200x200x400 mm "
                               "walled box, FP64, dt=1e-6 s, Kn=1e4, e=0.5, mu=0.3, cell=2*r_circ,max",
                   "particles": sc.n_particles, "candidates_per_step": stats["n_pairs"],
                   "gjk_pairs_per_step": stats["n_sphere_pass"], "contacts_per_step": stats["n_contacts"],
                   # pair tests: candidate pairs through GJK (the bounding-sphere
                   # test decides the others without GJK)
                   "pair_tests_per_s": stats["n_sphere_pass"] * args.steps / wall,
                   "candidates_per_s": stats["n_pairs"] * args.steps / wall,
                   "stage_ms_per_step": {k: round(v, 5) for k, v in main_leg["stage_ms"].items()},
                   "tuning": None if main_leg["tuned"] is None else main_leg["tuned"]["tuning"],
                   "untuned_ms_per_step_eager": None if main_leg["tuned"] is None
                   else main_leg["tuned"]["default_ms_per_step"],
                   "tuned_ms_per_step_eager": None if main_leg["tuned"] is None else main_leg["tuned"]["ms_per_step"]},
        "roofline": main_leg["roofline"],
        "roofline_gjk": roof_gjk,
        "cpu_baseline": cpu,
        "cpu_baseline_reference_structure": cpu_py,
        "contacting": contacting,
        "batched": batched,
        "batched_vs_plain": None if batched is None else batched["batched_vs_plain"],
        "differentiable": diff,
        "polydisperse": poly,
        "tangent_vs_forward": None if diff is None else diff["tangent_vs_forward"],
        "adjoint_vs_forward": None if diff is None else diff["adjoint_vs_forward"],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

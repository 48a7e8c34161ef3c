"""Autotuner of the launch choices of a Stepper.

The knobs are those of mwd_tuning (include/mwdiem_hip.h): how the cells are
built (key tables or a radix sort of the keys), GJK lanes per pair (so pairs
per wave), the GJK grid cap, the sphere-test block, the particles per block
of the fused force reduction + integrator, the occupancy asked of its
register allocation, whether its vertex pass skips padding slots and the
cells per block of the pair emission (the EPA grid cap is left to the
caller, see below). None of them changes a bit
of the results, so the tuner times them on the live state, restoring it
before every trial, and leaves the stepper on the fastest choice.

For scenes on the physical law the cell size is a free parameter as well:
any cell at least the largest particle extent finds every contact with the
full half-shell stencil. tune_cell_scale() times widened grids (fewer, fuller
cells; more candidate pairs). The contact set does not change with the cell
size; the order of the candidates, and with it the order in which each
particle's contact forces are summed, does, so results differ in rounding.
The reference law keeps the reference's grid (its 7-offset stencil and
candidate order depend on it), and tune_cell_scale() refuses such scenes.
"""
from __future__ import annotations

from dataclasses import replace
from typing import Sequence

import torch

from . import _lib
from .scene import PHYSICAL, SceneData, grid_for
from .stepper import Stepper

SPACE = {
    "verlet_block": (64, 128, 256),
    "gjk_group": (4, 8, 16),
    "gjk_blocks": (256, 1024, 4096),
    "sphere_block": (128, 256, 512),
    "cell_pass": (_lib.CELLS_TABLE, _lib.CELLS_SORTED),
    "cell_items": (0, 1, 2, 4),
    "verlet_waves": (1, 3, 4),
    "vertex_pack": (0, 1),
    "fill_cells": (64, 16, 4),
    "fill_split": (1, 4, 16),
    "gjk_buckets": (0, 1),
}
# Not searched: epa_blocks (the EPA grid cap) and epa_fused (EPA inside the
# GJK launch). Their best values follow the contact count, which grows over
# a run (a pour starts with none), so a choice timed on the first steps can
# starve the later ones; set them with Stepper.set_tuning when the contact
# count is known. The fused form saves the EPA launch (3 us on 100,000
# settling hulls, 12 us on a 4,096-hull pile) and costs about 3 % of the
# step at 14,000 contacts among 626,000 GJK survivors, where each wave runs
# the EPA of the contacts it found one after another.


def time_steps(S: Stepper, steps: int) -> float:
    """Milliseconds per eager step over `steps` steps on the stepper's stream."""
    S._enter()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(S.stream):
        e0.record()
    S.step(steps)
    with torch.cuda.stream(S.stream):
        e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) / steps


def autotune(S: Stepper, steps: int = 10, passes: int = 1, space: dict | None = None) -> dict:
    """Coordinate descent over the knobs: each knob in turn takes the value
    with the shortest step (the others held), `passes` times over. Every trial
    starts from the state the stepper had on entry, which it has again on
    return. Returns {"tuning", "ms_per_step", "default_ms_per_step", "trials"}."""
    space = dict(SPACE if space is None else space)
    snap = S.snapshot()
    S.check()
    best = S.tuning()
    trials = []

    def trial(choice: dict) -> float:
        S.restore(snap)
        S.set_tuning(**choice)
        S.step(2)  # first launches of a new variant
        S.restore(snap)
        ms = time_steps(S, steps)
        S.check()
        trials.append((dict(choice), ms))
        return ms

    start_ms = trial(best)
    best_ms = start_ms
    for _ in range(passes):
        for knob, values in space.items():
            for v in values:
                if v == best[knob]:
                    continue
                cand = dict(best, **{knob: v})
                try:
                    ms = trial(cand)
                except _lib.MwdError:  # a choice this workspace cannot take (no key tables)
                    continue
                if ms < best_ms:
                    best, best_ms = cand, ms
    S.restore(snap)
    S.set_tuning(**best)
    S.stream.synchronize()
    return {"tuning": best, "ms_per_step": best_ms, "default_ms_per_step": start_ms, "trials": trials}


def regrid(sc: SceneData, scale: float) -> SceneData:
    """The scene on a grid of cells `scale` times its particle extent."""
    cell, ncx, ncy, ncz, kmin, kspan = grid_for(sc.box, scale * sc.max_particlesize)
    return replace(sc.copy(), cell=cell, ncx=ncx, ncy=ncy, ncz=ncz, kmin=kmin, kspan=kspan)


def tune_cell_scale(scenes: Sequence[SceneData], scales=(1.0, 1.25, 1.5), steps: int = 10, **stepper_kw):
    """Time the scenes on each cell scale (one Stepper per scale, default
    launch choices, `steps` eager steps from the initial state) and return
    (best_scale, {scale: ms_per_step}). Physical-law scenes only."""
    if any(sc.law_p.mode != PHYSICAL or sc.law_w.mode != PHYSICAL for sc in scenes):
        raise ValueError("the cell size is free on the physical law only")
    ms = {}
    for f in scales:
        S = Stepper([regrid(sc, f) for sc in scenes], **stepper_kw)
        S.step(2)
        ms[f] = time_steps(S, steps)
        S.check()
        S.close()
        del S
        torch.cuda.empty_cache()
    best = min(ms, key=ms.get)
    return best, ms

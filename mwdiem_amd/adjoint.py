"""Reverse-mode gradients of bed outcomes, with step checkpointing.

The forward run keeps a device copy of the state every `every` steps. The
reverse sweep takes the segments last to first: it restores a segment's
checkpoint, recomputes the segment's steps keeping each step's input state,
then reverses those steps one at a time with mwd_step_adjoint -- the
hand-written adjoint kernels of the Verlet step, the physical contact law
and EPA's final-face geometry. The recomputed steps push their input states
and contact-stage outputs onto a device tape (mwd_step_record) and the
adjoint steps pop them (mwd_step_adjoint_taped), so the topology is the
forward's and the contact stages run once per recomputed step; without the
tape the input states are copied aside and the contact stages rerun on the
restored state. Memory: the checkpoints plus one segment of tape slots. The result equals the forward-mode tangents'
(mwdiem_amd.diff) to rounding; the adjoint costs one recomputed forward
step and one reverse step per step, whatever the number of parameters.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .diff import _with_theta, bed_outcome
from .stepper import STATE_KEYS, Stepper

SNAP_KEYS = STATE_KEYS + ("radius", "wall_c")


def _set_init(S: Stepper, init):
    if isinstance(init, dict):
        S.set_state(init)
    elif init is not None:
        S.set_state({"v": init, "v12": init})


def adjoint_run(sc, init, theta, steps: int, every: int = 100, top_frac: float = 0.1, seed=None,
                steps_per_graph: int = 0, timings: dict | None = None, tuning: dict | None = None,
                tape: bool = True):
    """Run `steps` steps of the scene with law parameters theta = (Kn, en,
    mu) from `init` (None, initial velocities or a state dict) and return
    (outcome, d outcome/d theta) by the adjoint. With steps_per_graph the
    forward and recomputed segments replay captured graphs. With tape (the
    default) the recomputed steps keep their input states and contact-stage
    outputs on the device (one slot per step of a segment) and the adjoint
    steps take them back instead of rerunning the contact stages; with
    steps_per_graph a segment is then one graph of recorded steps and one of
    adjoint steps. Without the tape, each step's input state is saved and put
    back by copy launches and each adjoint step reruns the contact stages. The
    gradient is the same bit for bit either way. timings, if given, receives
    the wall times (s) of the forward run, of the outcome with its gradient
    (the sweep's seed), and of the reverse sweep; tuning: launch choices."""
    import time
    S = Stepper(_with_theta(sc, theta))
    if tuning:
        S.set_tuning(**tuning)
    _set_init(S, init)
    S.set_law_seeds(seed)
    W = sc.n_walls
    t0 = time.perf_counter()
    if steps_per_graph:
        S.capture(steps_per_graph)
    ckpt = []
    for n0 in range(0, steps, every):
        ckpt.append(S.snapshot())
        S.run(min(every, steps - n0))
    S.check()
    torch.cuda.synchronize(S.device)
    t1 = time.perf_counter()
    with torch.enable_grad():
        com = S.t["com"].detach().clone().requires_grad_(True)
        L = bed_outcome(com, W, top_frac)
        (g,) = torch.autograd.grad(L, com)
    lam = {k: torch.zeros_like(S.t[k]) for k in _lib.TAN_ARRAYS}
    lam["com"].copy_(g)
    ptrs = (C.c_void_p * len(_lib.TAN_ARRAYS))(*[lam[k].data_ptr() for k in _lib.TAN_ARRAYS])
    dth = torch.zeros(_lib.N_TAN, dtype=torch.float64, device=S.device)
    # the outcome and its own gradient (torch; its first call in a process
    # also loads torch's kernels) are timed apart from the sweep
    torch.cuda.synchronize(S.device)
    t1s = time.perf_counter()
    stream = S.stream.cuda_stream
    if tape:
        # each recorded step pushes its input state and contact-stage outputs
        # onto the device tape and each taped adjoint step pops them; with
        # steps_per_graph both directions replay graphs of that many steps
        # (captured once per length, on the stepper's fixed buffers)
        S.tape_init(min(every, steps))
        g_rec, g_adj = {}, {}

        def replay(n, cache, make):
            while n > 0:
                k = min(steps_per_graph, n)
                if k not in cache:
                    cache[k] = make(k)
                with torch.cuda.stream(S.stream):
                    cache[k].replay()
                n -= k
        for j in reversed(range(len(ckpt))):
            n = min(every, steps - j * every)
            S.restore(ckpt[j])
            S.tape_rewind()
            if steps_per_graph:
                replay(n, g_rec, S.capture_record)
                replay(n, g_adj, lambda k: S.capture_adjoint(ptrs, dth, taped=True, steps=k))
            else:
                S.step_record(n)
                for _ in range(n):
                    S.step_adjoint_taped(ptrs, dth)
        del g_rec, g_adj
    else:
        # one segment's step inputs, saved and put back with one copy launch
        # each; with steps_per_graph the recomputed step and the adjoint step
        # replay graphs of one step each
        live = [S.t[k] for k in SNAP_KEYS]
        slots = [[torch.empty_like(t) for t in live] for _ in range(min(every, steps))]
        g_step = g_adj1 = None
        for j in reversed(range(len(ckpt))):
            n0 = j * every
            n1 = min(n0 + every, steps)
            S.restore(ckpt[j])
            for i in range(n1 - n0):
                _lib.copy_arrays(slots[i], live, stream)
                if steps_per_graph:
                    if g_step is None:
                        g_step = S.capture(1)
                    with torch.cuda.stream(S.stream):
                        g_step.replay()
                else:
                    S.step(1)
            for i in reversed(range(n1 - n0)):
                _lib.copy_arrays(live, slots[i], stream)
                if g_adj1 is not None:
                    with torch.cuda.stream(S.stream):
                        g_adj1.replay()
                else:
                    S.step_adjoint(ptrs, dth)
                    if steps_per_graph:
                        g_adj1 = S.capture_adjoint(ptrs, dth)
        del slots
    S.check()
    torch.cuda.synchronize(S.device)
    t2 = time.perf_counter()
    if timings is not None:
        timings["forward_s"] = t1 - t0
        timings["outcome_s"] = t1s - t1
        timings["reverse_s"] = t2 - t1s
    out = float(L.detach().cpu()), dth.cpu().numpy()
    S.close()
    return out


class BedOutcomeAdjoint(torch.autograd.Function):
    """outcome = BedOutcomeAdjoint.apply(theta, scene, init, steps, every):
    theta a float64 tensor (Kn, en, mu); backward returns grad_output * the
    adjoint gradient."""

    @staticmethod
    def forward(ctx, theta, sc, init, steps, every=100):
        L, g = adjoint_run(sc, init, theta.detach().cpu().numpy(), int(steps), int(every))
        ctx.save_for_backward(torch.from_numpy(np.asarray(g)).to(theta.dtype))
        return theta.new_tensor(L)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g.to(grad_out.device), None, None, None, None

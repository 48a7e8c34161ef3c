"""Gradients of bed outcomes with respect to the contact law's parameters.

The step is differentiated in forward (tangent) mode: every state array
carries its derivatives with respect to three law parameters theta =
(Kn, en, mu), advanced by the same kernels in dual-number arithmetic
(mwd_step_tangent), so the values are the plain step's bit for bit. With
three parameters and a scalar outcome this gives the same gradient as a
reverse-mode adjoint, needs no checkpoints and costs one tangent pass.
Contact topology is held fixed within each step: which pairs touch and EPA's
final face are taken from the values; depth, normal and contact points are
differentiated through that face's support vertices.

BedOutcome wraps a run as a torch.autograd.Function of theta.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .scene import Law4, SceneData
from .stepper import Stepper


def bed_outcome(com: torch.Tensor, n_walls: int, top_frac: float = 0.1) -> torch.Tensor:
    """Mean height of the top `top_frac` of particles plus the centroid spread
    (root-mean-square distance of the particle centres from their mean)."""
    c = com[n_walls:]
    k = max(1, int(math.ceil(top_frac * c.shape[0])))
    top = torch.topk(c[:, 2], k).values
    spread = torch.sqrt(((c - c.mean(dim=0)) ** 2).sum(dim=1).mean())
    return top.mean() + spread


def outcome_and_grad(S: Stepper, top_frac: float = 0.1):
    """Outcome of the stepper's current state (first scene) and its gradient
    with respect to theta, from the carried tangents of com."""
    torch.cuda.synchronize(S.device)
    W = S.scenes[0].n_walls
    with torch.enable_grad():  # also when called from an autograd.Function's forward
        com = S.t["com"].detach().clone().requires_grad_(True)
        L = bed_outcome(com, W, top_frac)
        (g,) = torch.autograd.grad(L, com)
    dL = torch.einsum("kij,ij->k", S.tan["com"], g)
    return float(L.detach().cpu()), dL.detach().cpu().numpy()


def _with_theta(sc: SceneData, theta) -> SceneData:
    Kn, en, mu = (float(x) for x in theta)
    out = sc.copy()
    out.law_p = Law4(Kn=Kn, en=en, et=sc.law_p.et, mu=mu, mode=sc.law_p.mode)
    out.law_w = Law4(Kn=Kn, en=en, et=sc.law_w.et, mu=mu, mode=sc.law_w.mode)
    return out


def form_pile(sc: SceneData, v0: np.ndarray | None = None, steps: int = 20_000, dt: float | None = None,
              steps_per_graph: int = 50, spawn=None, tuning: dict | None = None) -> dict:
    """Let the scene's particles fall and settle into a pile (its own law;
    dt, if given, replaces the scene's time step) and return the state (numpy
    arrays of the state keys) to start runs from. With spawn (scene.Spawn,
    e.g. from scene.config_pour) the particles enter on that schedule: the
    pile is poured. tuning: launch choices (Stepper.set_tuning; results do
    not depend on them)."""
    f = sc.copy()
    if dt is not None:
        f.dt = dt
    S = Stepper(f, spawn=spawn)
    if tuning:
        S.set_tuning(**tuning)
    if v0 is not None:
        S.set_state({"v": v0, "v12": v0})
    S.capture(steps_per_graph)
    S.run(steps)
    S.check()
    st = S.numpy_state()
    S.close()
    return st


def run_outcome(sc: SceneData, init, theta, steps: int, tangents: bool = True,
                steps_per_graph: int = 0, top_frac: float = 0.1, tuning: dict | None = None):
    """Run `steps` steps of the scene with law parameters theta = (Kn, en, mu)
    (both laws) from `init` -- None (the scene's initial state), an array of
    initial velocities, or a state dict; return (outcome, d outcome/d theta)
    (gradient None without tangents). tuning: launch choices."""
    S = Stepper(_with_theta(sc, theta))
    if tuning:
        S.set_tuning(**tuning)
    if isinstance(init, dict):
        S.set_state(init)
    elif init is not None:
        S.set_state({"v": init, "v12": init})
    if tangents:
        S.enable_tangents()
    if steps_per_graph:
        S.capture(steps_per_graph, tangent=tangents)
        S.run(steps)
    elif tangents:
        S.step_tangent(steps)
    else:
        S.step(steps)
    S.check()
    if tangents:
        L, g = outcome_and_grad(S, top_frac)
    else:
        torch.cuda.synchronize(S.device)
        L, g = float(bed_outcome(S.t["com"], sc.n_walls, top_frac).cpu()), None
    S.close()
    return L, g


class BedOutcome(torch.autograd.Function):
    """outcome = BedOutcome.apply(theta, scene, init, steps): theta a float64
    tensor (Kn, en, mu), init as for run_outcome; backward returns
    grad_output * d outcome/d theta."""

    @staticmethod
    def forward(ctx, theta, sc, init, steps, steps_per_graph=0):
        L, g = run_outcome(sc, init, theta.detach().cpu().numpy(), int(steps), True, int(steps_per_graph))
        ctx.save_for_backward(torch.from_numpy(g).to(theta.dtype))
        return theta.new_tensor(L)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g.to(grad_out.device), None, None, None, None

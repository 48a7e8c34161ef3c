"""Tangent (forward-mode) step on the GPU: values identical to the plain step,
derivatives against central finite differences of the plain step."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pile(n=256, seed=5):
    """A poured pile fast enough that hulls collide within tens of steps."""
    from mwdiem_amd import scene
    return scene.config_pile(seed=seed, n=n, v_pour=3000.0)


def test_tangent_step_values_are_bitwise(gpu, dense_phys):
    """mwd_step_tangent computes the plain step's values bit for bit."""
    from mwdiem_amd.stepper import Stepper
    A = Stepper(dense_phys)
    B = Stepper(dense_phys)
    B.enable_tangents()
    A.step(40)
    B.step_tangent(40)
    A.check()
    B.check()
    a, b = A.numpy_state(), B.numpy_state()
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    # something moved and some tangent is nonzero
    assert np.abs(B.numpy_tangents()["com"]).max() > 0.0


def test_tangent_graph_replay_equals_eager(gpu, dense_phys):
    from mwdiem_amd.stepper import Stepper
    A = Stepper(dense_phys)
    A.enable_tangents()
    A.step_tangent(30)
    B = Stepper(dense_phys)
    B.enable_tangents()
    B.capture(10, tangent=True)
    B.run(30)
    ta, tb = A.numpy_tangents(), B.numpy_tangents()  # both streams finished
    for k in ta:
        assert np.array_equal(ta[k], tb[k]), k
    a, b = A.numpy_state(), B.numpy_state()
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def _com_after(sc, v0, theta, steps):
    from mwdiem_amd.diff import _with_theta
    from mwdiem_amd.stepper import Stepper
    S = Stepper(_with_theta(sc, theta))
    S.set_state({"v": v0, "v12": v0})
    S.step(steps)
    S.check()
    out = S.numpy_state()["com"]
    S.close()
    return out


@pytest.mark.parametrize("k", [0, 1, 2])
def test_state_tangent_matches_central_differences(gpu, k):
    """d com / d theta_k after 400 steps of a poured pile against central
    differences (h = 1e-6 relative) of the plain step: every particle agrees
    to 1e-4 of the largest derivative (differences of order h^2 and the
    rounding of the differenced positions bound it; contact topology is the
    same in the three runs over so short a run, in which the first contacts
    form)."""
    from mwdiem_amd.diff import _with_theta
    from mwdiem_amd.stepper import Stepper
    sc, v0 = _pile()
    theta = np.array([1e4, 0.5, 0.3])
    steps = 400
    S = Stepper(_with_theta(sc, theta))
    S.set_state({"v": v0, "v12": v0})
    S.enable_tangents()
    S.step_tangent(steps)
    S.check()
    tan = S.numpy_tangents()["com"][k]
    h = 1e-6 * theta[k]
    tp, tm = theta.copy(), theta.copy()
    tp[k] += h
    tm[k] -= h
    fd = (_com_after(sc, v0, tp, steps) - _com_after(sc, v0, tm, steps)) / (2 * h)
    W = sc.n_walls
    scale = np.abs(fd[W:]).max()
    print(f"theta_{k}: max|tangent| {np.abs(tan[W:]).max():.6e} max|fd| {scale:.6e} "
          f"max|diff| {np.abs(tan[W:] - fd[W:]).max():.3e}")
    assert scale > 0
    assert np.abs(tan[W:] - fd[W:]).max() <= 1e-4 * scale


def test_bed_outcome_gradient_against_central_differences(gpu):
    """Gradient of the bed outcome (mean height of the top 10% + centroid
    spread) after 200 steps against central differences with h = 1e-4
    relative: each component within 1e-5 relative (measured: about 1e-7)."""
    import torch
    from mwdiem_amd.diff import BedOutcome, run_outcome
    sc, v0 = _pile()
    theta = np.array([1e4, 0.5, 0.3])
    steps = 200
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    L = BedOutcome.apply(th, sc, v0, steps)
    L.backward()
    g = th.grad.numpy()
    L0, g0 = run_outcome(sc, v0, theta, steps)
    assert float(L) == L0 and np.array_equal(g, g0)
    fd = np.zeros(3)
    for k in range(3):
        h = 1e-4 * theta[k]
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        Lp, _ = run_outcome(sc, v0, tp, steps, tangents=False)
        Lm, _ = run_outcome(sc, v0, tm, steps, tangents=False)
        fd[k] = (Lp - Lm) / (2 * h)
    rel = np.abs(g - fd) / np.abs(fd)
    print(f"outcome {L0:.12f} grad {g} fd {fd} rel {rel}")
    assert np.all(np.abs(fd) > 0) and np.all(rel <= 1e-5)


def test_bed_outcome_gradient_against_torch_cpu_reference(gpu):
    """The GPU's forward-mode gradient of the bed outcome after 400 steps of
    the 3 m/s pour (about 2,000 contacts) against the reverse-mode gradient of
    the FP64 PyTorch-CPU restatement of the step (oracle/torch_step.py, same
    topology per step): outcome within 1e-12, each gradient component within
    1e-6 relative."""
    import torch
    from mwdiem_amd.diff import bed_outcome, run_outcome
    from oracle import oracle as O
    from oracle import torch_step as TS
    sc, v0 = _pile()
    theta = np.array([1e4, 0.5, 0.3])
    steps = 400
    L0, g0 = run_outcome(sc, v0, theta, steps)
    init = O.initial_state(sc)
    init["v"] = v0.copy()
    init["v12"] = v0.copy()
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    st, counts = TS.run(sc, init, th, steps)
    L = bed_outcome(st["com"], sc.n_walls)
    L.backward()
    g = th.grad.numpy()
    rel = np.abs(g0 - g) / np.abs(g)
    print(f"contacts {sum(counts)} outcome {L0:.15f} vs {float(L):.15f} grad {g0} vs {g} rel {rel}")
    assert sum(counts) > 500
    assert abs(L0 - float(L)) <= 1e-12 * abs(L0)
    assert np.all(rel <= 1e-6)

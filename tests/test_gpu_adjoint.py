"""Reverse-mode adjoint step (hand-written adjoint kernels) on the GPU: the
dot-product identity against the tangent step on one step, and the bed
outcome's gradient against the forward-mode tangents over a run."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _pile(n=256, seed=5):
    from mwdiem_amd import scene
    return scene.config_pile(seed=seed, n=n, v_pour=3000.0)


def _state_with_contacts(steps=420):
    """The 3 m/s pour after its first impacts (contacts in progress)."""
    from mwdiem_amd.stepper import Stepper
    sc, v0 = _pile()
    S = Stepper(sc)
    S.set_state({"v": v0, "v12": v0})
    S.step(steps)
    S.check()
    st = S.numpy_state()
    S.close()
    return sc, st


def test_adjoint_step_satisfies_the_dot_product_identity(gpu):
    """For one step J (state and theta): lam . (J dx) == (J^T lam) . dx with dx
    a random state direction (moving rows) or a parameter direction; lam is
    random. Both sides within 1e-10 relative."""
    from mwdiem_amd import _lib
    from mwdiem_amd.stepper import Stepper
    sc, st = _state_with_contacts()
    W, N = sc.n_walls, sc.n_rows
    rng = np.random.Generator(np.random.PCG64(17))
    keys = _lib.TAN_ARRAYS
    # tangent side: direction 0 a random state direction, 1 and 2 none; law seeds for (Kn, en, mu)
    T = Stepper(sc)
    T.set_state(st)
    T.enable_tangents()
    dx = {}
    for k in keys:
        d = np.zeros_like(st[k])
        d[W:] = rng.normal(size=d[W:].shape)
        if k == "verts":
            live = np.arange(sc.vmax)[None, :] < sc.nverts[:, None]
            d[~live] = 0.0
        dx[k] = d
        T.tan[k][0].copy_(torch.from_numpy(d))
    torch.cuda.synchronize()
    T.step_tangent(1)
    T.check()
    tan = T.numpy_tangents()
    assert T.status()["n_contacts"] > 5
    lam = {k: rng.normal(size=st[k].shape) for k in keys}
    for k in keys:
        lam[k][:W] = 0.0
    lhs_state = sum(float(np.sum(lam[k] * tan[k][0])) for k in keys)
    lhs_theta = np.array([sum(float(np.sum(lam[k] * tan[k][j])) for k in keys) for j in (1, 2)])
    # the same step with seeds only (no state direction) gives the theta part of direction 0 too
    T2 = Stepper(sc)
    T2.set_state(st)
    T2.enable_tangents()
    T2.step_tangent(1)
    tan2 = T2.numpy_tangents()
    th_tan = np.array([sum(float(np.sum(lam[k] * tan2[k][j])) for k in keys) for j in range(3)])
    lhs_state -= th_tan[0]  # direction 0 carried theta_0 as well
    assert np.allclose(lhs_theta, th_tan[1:], rtol=1e-12)
    # adjoint side
    A = Stepper(sc)
    A.set_state(st)
    A.set_law_seeds()
    lt = {k: torch.from_numpy(lam[k].copy()).to(A.device) for k in keys}
    ptrs = (C.c_void_p * len(keys))(*[lt[k].data_ptr() for k in keys])
    dth = torch.zeros(3, dtype=torch.float64, device=A.device)
    torch.cuda.synchronize()
    A.step_adjoint(ptrs, dth)
    A.check()
    torch.cuda.synchronize()
    lam_n = {k: lt[k].cpu().numpy() for k in keys}
    rhs_state = sum(float(np.sum(lam_n[k] * dx[k])) for k in keys)
    rhs_theta = dth.cpu().numpy()
    print(f"state: {lhs_state:.15e} vs {rhs_state:.15e}; theta: {th_tan} vs {rhs_theta}")
    assert abs(lhs_state - rhs_state) <= 1e-10 * abs(lhs_state)
    assert np.all(np.abs(th_tan - rhs_theta) <= 1e-10 * np.abs(th_tan) + 1e-300)


def test_adjoint_gradient_matches_tangents(gpu):
    """The bed outcome's gradient after 400 steps of the 3 m/s pour (about
    2,000 contacts): adjoint with checkpoints every 64 steps against the
    forward-mode tangents, within 1e-9 relative."""
    from mwdiem_amd.adjoint import BedOutcomeAdjoint, adjoint_run
    from mwdiem_amd.diff import run_outcome
    sc, v0 = _pile()
    theta = np.array([1e4, 0.5, 0.3])
    L_t, g_t = run_outcome(sc, v0, theta, 400)
    L_a, g_a = adjoint_run(sc, v0, theta, 400, every=64)
    rel = np.abs(g_a - g_t) / np.abs(g_t)
    print(f"outcome {L_a} vs {L_t}; adjoint {g_a} tangents {g_t} rel {rel}")
    assert L_a == L_t
    assert np.all(rel <= 1e-9)
    # the same sweep with the forward, recomputed and adjoint steps replayed
    # from graphs (fused state copies between the replays): bit for bit
    L_g, g_g = adjoint_run(sc, v0, theta, 400, every=64, steps_per_graph=16)
    assert L_g == L_a and np.array_equal(g_g, g_a), (g_g, g_a)
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    out = BedOutcomeAdjoint.apply(th, sc, v0, 400, 128)
    out.backward()
    assert float(out) == L_t and np.allclose(th.grad.numpy(), g_a, rtol=1e-12, atol=0)


def test_taped_sweep_equals_the_rerun_sweep(gpu):
    """The reverse sweep whose recomputed steps keep their contact-stage
    outputs on the tape gives the gradient of the sweep that reruns the
    contact stages in each adjoint step, bit for bit: eager, and with every
    step replayed from graphs; 200 steps of the 3 m/s pour from its first
    impacts, checkpoints every 64 steps (a short last segment)."""
    from mwdiem_amd.adjoint import adjoint_run
    sc, st = _state_with_contacts()
    theta = np.array([1e4, 0.5, 0.3])
    L0, g0 = adjoint_run(sc, st, theta, 200, every=64, tape=False)
    L1, g1 = adjoint_run(sc, st, theta, 200, every=64, tape=True)
    print(f"outcome {L0}; rerun {g0} taped {g1}")
    assert np.all(np.isfinite(g0)) and np.any(g0 != 0)
    assert L1 == L0 and np.array_equal(g1, g0), (g1, g0)
    L2, g2 = adjoint_run(sc, st, theta, 200, every=64, steps_per_graph=8, tape=True)
    assert L2 == L0 and np.array_equal(g2, g0), (g2, g0)


def test_tape_full_and_empty_raise(gpu):
    """A recorded step on a full tape and a taped adjoint step on an empty one
    copy nothing and raise the tape flag, which check() reports; the tape's
    slot count follows pushes and pops."""
    from mwdiem_amd import _lib
    from mwdiem_amd.stepper import Stepper
    sc, st = _state_with_contacts()
    S = Stepper(sc)
    S.set_state(st)
    S.set_law_seeds()
    assert S.tape_init(2) > 0
    S.step_record(2)
    S.check()
    S.step_record(1)
    with pytest.raises(_lib.MwdError, match="tape"):
        S.check()
    _lib.check(_lib.load().mwd_clear_flags(S.ws), "mwd_clear_flags")
    lt = {k: torch.zeros_like(S.t[k]) for k in _lib.TAN_ARRAYS}
    ptrs = (C.c_void_p * len(lt))(*[lt[k].data_ptr() for k in _lib.TAN_ARRAYS])
    dth = torch.zeros(3, dtype=torch.float64, device=S.device)
    S.step_adjoint_taped(ptrs, dth)
    S.step_adjoint_taped(ptrs, dth)
    S.check()
    S.step_adjoint_taped(ptrs, dth)
    with pytest.raises(_lib.MwdError, match="tape"):
        S.check()
    S.close()


def test_adjoint_refuses_the_reference_law(gpu, dense_scene):
    from mwdiem_amd import _lib
    from mwdiem_amd.stepper import Stepper
    S = Stepper(dense_scene)
    S.set_law_seeds()
    lt = {k: torch.zeros_like(S.t[k]) for k in _lib.TAN_ARRAYS}
    ptrs = (C.c_void_p * len(lt))(*[lt[k].data_ptr() for k in _lib.TAN_ARRAYS])
    with pytest.raises(_lib.MwdError):
        S.step_adjoint(ptrs, torch.zeros(3, dtype=torch.float64, device=S.device))

"""The FP64 PyTorch-CPU reference of the differentiable step (oracle/
torch_step.py), pinned on the CPU: its forward values follow the oracle's
step, and its reverse-mode gradient of the bed outcome matches central
differences of the oracle."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as O
from oracle import torch_step as TS

STEPS = 400  # a pour at 3 m/s: the first impacts, some 2,000 contacts in all
THETA = np.array([1e4, 0.5, 0.3])


@pytest.fixture(scope="module")
def pile():
    from mwdiem_amd import scene
    sc, v0 = scene.config_pile(seed=5, n=256, v_pour=3000.0)
    init = O.initial_state(sc)
    init["v"] = v0.copy()
    init["v12"] = v0.copy()
    return sc, init


def _oracle_run(sc, init, theta, steps):
    from mwdiem_amd.diff import _with_theta
    st = {k: v.copy() for k, v in init.items()}
    scx = _with_theta(sc, theta)
    n = 0
    for _ in range(steps):
        _, coll = O.step(st, scx)
        n += int((coll == 1).sum())
    return st, n


@pytest.fixture(scope="module")
def torch_run(pile):
    from mwdiem_amd.diff import bed_outcome
    sc, init = pile
    th = torch.tensor(THETA, dtype=torch.float64, requires_grad=True)
    st, counts = TS.run(sc, init, th, STEPS)
    L = bed_outcome(st["com"], sc.n_walls)
    L.backward()
    return st, counts, float(L.detach()), th.grad.numpy().copy()


def test_forward_follows_the_oracle(pile, torch_run):
    sc, init = pile
    st, counts, _, _ = torch_run
    ref, n = _oracle_run(sc, init, THETA, STEPS)
    assert n == sum(counts) and n > 500
    W = sc.n_walls
    for k in ("com", "q", "qall", "v", "w", "verts"):
        a, b = st[k].detach().numpy()[W:], ref[k][W:]
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max(), k


def test_gradient_matches_oracle_central_differences(pile, torch_run):
    """h = 1e-6 relative: each component within 1e-5 (measured about 2e-6)."""
    from mwdiem_amd.diff import bed_outcome
    sc, init = pile
    _, _, L, g = torch_run
    st0, _ = _oracle_run(sc, init, THETA, STEPS)
    assert abs(float(bed_outcome(torch.from_numpy(st0["com"]), sc.n_walls)) - L) <= 1e-12 * abs(L)
    fd = np.zeros(3)
    for k in range(3):
        h = 1e-6 * THETA[k]
        tp, tm = THETA.copy(), THETA.copy()
        tp[k] += h
        tm[k] -= h
        Lp = float(bed_outcome(torch.from_numpy(_oracle_run(sc, init, tp, STEPS)[0]["com"]), sc.n_walls))
        Lm = float(bed_outcome(torch.from_numpy(_oracle_run(sc, init, tm, STEPS)[0]["com"]), sc.n_walls))
        fd[k] = (Lp - Lm) / (2 * h)
    rel = np.abs(g - fd) / np.abs(fd)
    print(f"torch grad {g} oracle fd {fd} rel {rel}")
    assert np.all(np.abs(fd) > 0) and np.all(rel <= 1e-5)


def test_face_geometry_closed_form():
    """Two unit cubes overlapping by 0.25 along x: the face of the Minkowski
    difference nearest the origin is x = 0.25 away; depth 0.25, normal +-x,
    and the contact points lie in the overlap slab."""
    c = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], float)
    A = c
    B = c + np.array([0.75, 0.1, 0.2])
    col, depth, n, cpA, cpB, _ = O.gjk_epa(A, B, A.mean(0), B.mean(0))
    assert col == 1 and abs(depth - 0.25) < 1e-12
    # the same face through the torch restatement, from the oracle's support indices
    sc = type("S", (), {"vmax": 8, "nverts": np.array([8, 8], np.int32)})()
    st = {"com": np.stack([A.mean(0), B.mean(0)]), "verts": np.stack([A, B])}
    coll, face = O.contact_faces(np.array([[0, 1]], np.int32), st, sc)
    assert coll[0] == 1
    VA = torch.from_numpy(A[face[0, :3]])[None]
    VB = torch.from_numpy(B[face[0, 3:]])[None]
    d, nt, ca, cb = TS.face_geometry(VA, VB)
    assert abs(float(d[0]) - depth) < 1e-12
    assert np.allclose(np.abs(nt[0].numpy()), [1, 0, 0], atol=1e-12)
    assert np.allclose(ca[0].numpy(), cpA, atol=1e-12) and np.allclose(cb[0].numpy(), cpB, atol=1e-12)

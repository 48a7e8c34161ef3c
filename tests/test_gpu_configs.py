"""The workload configurations run on the GPU against the oracle:

  C1  plain small scene: 200 hulls of 12 points in a 10x10x30 mm box, the
      reference law, 1,000 steps from rest -- the candidate list identical to
      the oracle's at every step and the state within 1e-8 at the end; and the
      same scene agitated (contacts in progress) for 100 steps on both laws;
  C4  the batched form: all 256 hopper scenes (500-20,000 hulls each, per
      scene friction, stiffness and restitution) thrown about at 5 m/s rms and
      stepped for 250 steps in one
      Stepper, bit-identical to each scene stepped alone, and the largest
      scene and three others against the oracle;
  C5  the differentiable form at full size: the 4,096-hull pile poured on
      the device (scene.config_pour: 34 batches of 121 hulls entering the top
      of a 40x40x60 mm box, 102,662 steps), then the bed outcome's gradient
      with respect to (Kn, en, mu) over 20 steps of the settled pile (about
      154,000 contact events) by the adjoint with checkpoints every 10 steps,
      against the FP64 PyTorch-CPU reverse-mode reference
      (oracle/torch_step.py) on the same state; and the same for the dilute
      start of a fake pour (scene.config_pile) over 700 steps.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.helpers import max_rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1():
    from mwdiem_amd import scene
    return scene.config_small()


def _agitate(sc, st, seed, vrms):
    rng = np.random.Generator(np.random.PCG64(seed))
    W = sc.n_walls
    v = rng.normal(size=(sc.n_particles, 3)) * vrms
    st["v"][W:] = v
    st["v12"][W:] = v
    return st


def _lockstep(sc, st, steps, check_every=1):
    """Step the oracle and a Stepper side by side; the candidate lists must be
    identical at every checked step. Returns (GPU state, oracle state,
    contacts seen by the oracle)."""
    from mwdiem_amd.stepper import Stepper
    S = Stepper(sc)
    S.set_state({k: st[k] for k in O.STATE_KEYS})
    ref = {k: v.copy() for k, v in st.items()}
    n_c = 0
    for i in range(steps):
        pairs, coll = O.step(ref, sc)
        n_c += int((coll == 1).sum())
        S.step(1)
        if i % check_every == 0 or i == steps - 1:
            S.check()
            got = S.pairs()
            assert got.shape == pairs.shape and np.array_equal(got, pairs), f"step {i}"
    S.check()
    return S.numpy_state(), ref, n_c


def test_c1_small_scene_1000_steps(gpu, c1):
    sc = c1
    assert sc.n_particles == 200 and np.all(sc.nverts[sc.n_walls:] == 12)
    print("C1: 1000 steps from rest", flush=True)
    got, ref, n_c = _lockstep(sc, O.initial_state(sc), 1000)
    W = sc.n_walls
    for k in ("com", "q", "qall", "v", "verts"):
        assert max_rel(got[k][W:], ref[k][W:]) <= 1e-8, k
    assert np.all(got["v"][W:, 2] < 0)  # falling


def test_c1_small_scene_agitated_physical_law(gpu, c1):
    """C1 with random velocities of 1 m/s rms on the physical law: hulls
    collide within the run (some 3,500 contact events); candidate lists
    identical at every step and the state within 1e-8 after 100 steps."""
    from mwdiem_amd import scene
    sc = c1.copy()
    sc.law_p = scene.Law4(mode=scene.PHYSICAL)
    sc.law_w = scene.Law4(mode=scene.PHYSICAL)
    print("C1 agitated, physical law", flush=True)
    st = _agitate(sc, O.initial_state(sc), 31, 1000.0)
    got, ref, n_c = _lockstep(sc, st, 100)
    assert n_c > 1000
    W = sc.n_walls
    for k in ("com", "q", "v", "w"):
        assert max_rel(got[k][W:], ref[k][W:]) <= 1e-8, k


def test_c1_small_scene_agitated_reference_law_step_by_step(gpu, c1):
    """C1 with random velocities of 1 m/s rms on the reference law. Its
    damping takes mu for a mass (Q4) and is unstable at these sizes (Gn dt/m
    about 10): a contact multiplies velocities, and any rounding difference,
    about tenfold per step (1e8 mm/s within 5 steps, NaN within 15, on the
    oracle as on the GPU), so two runs cannot be compared after a few steps.
    Instead every step starts the oracle from the GPU's state: for 10 steps
    the candidate list is identical and positions and velocities are within
    1e-10; orientations, angular velocities and vertices are compared while
    the fastest hull turns by less than 0.1 rad per step (beyond that a
    last-ulp difference of a torque moves an orientation by that ulp times
    the turn angle, thousands of radians per step once the bed blows up).
    Only the normal force acts on the linear momentum (Q5), so positions and
    velocities stay comparable."""
    from mwdiem_amd.stepper import Stepper
    sc = c1
    st = _agitate(sc, O.initial_state(sc), 31, 1000.0)
    S = Stepper(sc)
    S.set_state({k: st[k] for k in O.STATE_KEYS})
    W = sc.n_walls
    compared_with_contacts = 0
    for i in range(10):
        cur = S.numpy_state()
        if not all(np.isfinite(cur[k]).all() for k in cur):
            break
        turn = float(np.linalg.norm(cur["w"][W:], axis=1).max()) * sc.dt
        ref = {k: v.copy() for k, v in cur.items()}
        pairs, coll = O.step(ref, sc)
        S.step(1)
        S.check()
        assert np.array_equal(S.pairs(), pairs), f"step {i}"
        got = S.numpy_state()
        for k in ("com", "v", "v12") + (("q", "w", "verts") if turn < 0.1 else ()):
            assert max_rel(got[k][W:], ref[k][W:]) <= 1e-10, (i, k)
        compared_with_contacts += int((coll == 1).any())
    assert compared_with_contacts >= 6


@pytest.fixture(scope="module")
def c4():
    from mwdiem_amd import scene
    scs = scene.config_batched(256)
    inits = []
    for s, sc in enumerate(scs):
        inits.append(_agitate(sc, O.initial_state(sc), 1000 + s, 5000.0))
    return scs, inits


C4_STEPS = 250


def test_c4_batched_equals_scenes_alone(gpu, c4):
    """All 256 scenes in one Stepper (one launch per stage over all of them)
    against each scene in a Stepper of its own: identical states bit for bit
    after 250 steps of hulls thrown about at 5 m/s rms, contacts in progress
    (NaN in the same places: in scene 192 one thin hull, whose rotational
    stability number sqrt(Kn r^2 / I_min) dt is 1.35 at that scene's Kn,
    spins up from step 75 and overflows at step 182, on the oracle as on the
    GPU -- the explicit integrator's limit at dt = 1e-6 s)."""
    from mwdiem_amd.stepper import Stepper
    scs, inits = c4
    print(f"C4: {len(scs)} scenes, {sum(sc.n_particles for sc in scs)} hulls", flush=True)
    B = Stepper(scs)
    B.set_state({k: np.concatenate([st[k] for st in inits]) for k in O.STATE_KEYS})
    n_contacts = 0
    for _ in range(C4_STEPS):
        B.step(1)
    s = B.check()
    n_contacts = s["n_contacts"]
    batched = B.numpy_state()
    rows = [B.scene_rows(i) for i in range(len(scs))]
    B.close()
    assert n_contacts > 10000, n_contacts
    for i, sc in enumerate(scs):
        S = Stepper(sc)
        S.set_state({k: inits[i][k] for k in O.STATE_KEYS})
        S.step(C4_STEPS)
        S.check()
        alone = S.numpy_state()
        S.close()
        for k in O.STATE_KEYS:
            assert np.array_equal(batched[k][rows[i]], alone[k], equal_nan=True), (i, k)


def test_c4_scenes_against_the_oracle(gpu, c4):
    """The largest scene (19,783 hulls) and three others: each Stepper's
    candidate list equals the oracle's every 10 steps, and the state after
    250 steps (some 6,000-66,000 contact events per scene) is within 1e-8 of
    the oracle's."""
    scs, inits = c4
    big = int(np.argmax([sc.n_particles for sc in scs]))
    for i in (big, 0, 77, 255):
        sc = scs[i]
        print(f"C4 scene {i}: {sc.n_particles} hulls", flush=True)
        got, ref, n_c = _lockstep(sc, {k: v.copy() for k, v in inits[i].items()}, C4_STEPS, check_every=10)
        assert n_c > 0, i
        W = sc.n_walls
        for k in ("com", "q", "v", "w"):
            assert max_rel(got[k][W:], ref[k][W:]) <= 1e-8, (i, k)


@pytest.fixture(scope="module")
def poured_pile():
    from mwdiem_amd import scene
    from mwdiem_amd.diff import form_pile
    sc, v_pour, spawn = scene.config_pour(n=4096)
    n_pour = spawn.every * ((sc.n_particles + spawn.per - 1) // spawn.per) + 40_000
    print(f"C5: pouring {n_pour} steps", flush=True)
    st0 = form_pile(sc, v_pour, steps=n_pour, spawn=spawn, steps_per_graph=20,
                    tuning=dict(fill_cells=4, gjk_group=4))
    return sc, st0


def test_c5_poured_pile_candidates_both_cell_passes(gpu, poured_pile):
    """The settled pile packs cells of up to about 40 hulls (the set-order
    walks past sixteen members, the key lists, the LDS and scratch tables):
    the candidate list of one step, order included, from the key-table pass
    and from the sorted pass (with the emission's block split) equals the
    oracle's."""
    from mwdiem_amd import _lib
    from mwdiem_amd.stepper import Stepper
    from oracle import oracle as O
    sc, st0 = poured_pile
    ref = O.candidates({k: v.copy() for k, v in st0.items()}, sc)
    for choice in (dict(cell_pass=_lib.CELLS_TABLE), dict(cell_pass=_lib.CELLS_SORTED, fill_cells=4, fill_split=16),
                   dict(cell_pass=_lib.CELLS_TABLE, cell_items=0, fill_cells=16, fill_split=4)):
        S = Stepper(sc)
        S.set_state(st0)
        S.set_tuning(**choice)
        S.step(1)
        S.check()
        got = S.pairs()
        S.close()
        print(f"C5 candidates {choice}: {len(got)} pairs (oracle {len(ref)})")
        assert np.array_equal(got, ref), choice


def test_c5_poured_pile_adjoint_against_torch_cpu_reference(gpu, poured_pile, steps=20):
    """The pile poured on the device (physical law, Kn=1e4, en=0.5, mu=0.3):
    d(bed outcome)/d(Kn, en, mu) over 20 steps of the settled pile (154,000
    contact events) by the adjoint (checkpoints every 10 steps, graph-
    replayed) against the PyTorch-CPU reverse-mode reference on the same
    state, within 1e-9 relative, and the outcome within 1e-12. (Over 100
    steps the two FP64 trajectories part by rounding: 8 of 767,000 contact
    events differ and the gradients differ by 2e-4 -- the pile is chaotic,
    so the horizon is kept where both follow one trajectory.)"""
    from mwdiem_amd.adjoint import adjoint_run
    from mwdiem_amd.diff import bed_outcome
    from oracle import torch_step as TS
    sc, st0 = poured_pile
    assert all(np.isfinite(a).all() for a in st0.values())
    W = sc.n_walls
    # settled: the median hull moves slower than 10 mm/s
    assert np.median(np.linalg.norm(st0["v"][W:], axis=1)) < 10.0
    theta = np.array([1e4, 0.5, 0.3])
    print("C5: adjoint", flush=True)
    L_a, g_a = adjoint_run(sc, st0, theta, steps, every=10, steps_per_graph=10,
                           tuning=dict(fill_cells=4, gjk_group=4))
    print("C5: torch-CPU reference", flush=True)
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    st, counts = TS.run(sc, {k: v.copy() for k, v in st0.items()}, th, steps)
    L = bed_outcome(st["com"], W)
    L.backward()
    g_t = th.grad.numpy()
    rel = np.abs(g_a - g_t) / np.abs(g_t)
    print(f"C5: contacts {sum(counts)}; outcome {L_a} vs {float(L)}; adjoint {g_a} torch {g_t} rel {rel}")
    assert sum(counts) > 7000 * steps
    assert abs(L_a - float(L)) <= 1e-12 * abs(L_a)
    assert np.all(rel <= 1e-9)


def test_c5_dilute_start_adjoint_against_torch_cpu_reference(gpu):
    """The dilute start (scene.config_pile: hulls on a lattice thrown at
    1 m/s rms, the first impacts): the adjoint after 700 steps against the
    PyTorch-CPU reference within 1e-9 relative."""
    from mwdiem_amd import scene
    from mwdiem_amd.adjoint import adjoint_run
    from mwdiem_amd.diff import bed_outcome
    from oracle import torch_step as TS
    sc, v0 = scene.config_pile(n=4096)
    theta = np.array([1e4, 0.5, 0.3])
    steps = 700
    print("C5: adjoint", flush=True)
    L_a, g_a = adjoint_run(sc, v0, theta, steps, every=100, steps_per_graph=20)
    print("C5: torch-CPU reference", flush=True)
    init = O.initial_state(sc)
    init["v"] = v0.copy()
    init["v12"] = v0.copy()
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    st, counts = TS.run(sc, init, th, steps)
    L = bed_outcome(st["com"], sc.n_walls)
    L.backward()
    g_t = th.grad.numpy()
    rel = np.abs(g_a - g_t) / np.abs(g_t)
    print(f"C5: contacts {sum(counts)}; outcome {L_a} vs {float(L)}; adjoint {g_a} torch {g_t} rel {rel}")
    assert sum(counts) > 500
    assert abs(L_a - float(L)) <= 1e-12 * abs(L_a)
    assert np.all(rel <= 1e-9)

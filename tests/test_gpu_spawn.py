"""Spawning on the device (MWDiem.py:685-714): particles join the run in
batches on a schedule the device counts, against the oracle stepping the same
active range, eagerly and from graph replays."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import max_rel

pytestmark = pytest.mark.gpu


def _pour(n=300):
    from mwdiem_amd import scene
    # a narrow box poured fast: batches of one layer every 300 steps
    return scene.config_pour(n=n, L=(12.0, 12.0, 14.0), v_pour=10000.0, every=300, size=(0.6, 0.9))


def test_spawned_rows_join_on_schedule_and_match_the_oracle(gpu):
    from mwdiem_amd.stepper import Stepper
    sc, v, sp = _pour()
    W = sc.n_walls
    st = O.initial_state(sc)
    st["v"][:] = v
    st["v12"][:] = v
    S = Stepper(sc, spawn=sp)
    S.set_state({"v": v, "v12": v})
    ref = {k: a.copy() for k, a in st.items()}
    steps, n_c = 2000, 0
    for i in range(steps):
        n_act = sp.active_at(i, sc.n_rows)
        pairs, coll = O.step(ref, sc, n_active=n_act)
        n_c += int((coll == 1).sum())
        S.step(1)
        if i % 50 == 0 or i == steps - 1:
            S.check()
            got = S.pairs()
            assert np.array_equal(got, pairs), f"step {i}"
    got = S.numpy_state()
    n_act = sp.active_at(steps - 1, sc.n_rows)
    print(f"spawned {n_act - W} of {sc.n_particles}; {n_c} contact events", flush=True)
    assert W + 6 * sp.per <= n_act < sc.n_rows and n_c > 0
    for k in ("com", "q", "v", "w"):
        assert max_rel(got[k][W:n_act], ref[k][W:n_act]) <= 1e-8, k
    # rows still waiting have not moved
    for k in ("com", "v", "verts"):
        assert np.array_equal(got[k][n_act:], st[k][n_act:]), k


def test_spawn_graph_replays_equal_eager_steps(gpu):
    from mwdiem_amd.stepper import Stepper
    sc, v, sp = _pour()
    out = []
    for graph in (False, True):
        S = Stepper(sc, spawn=sp)
        S.set_state({"v": v, "v12": v})
        if graph:
            S.capture(20)
        S.run(1000)
        S.check()
        out.append(S.numpy_state())
        S.close()
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k]), k


def test_spawn_refusals(gpu, small_scene):
    import ctypes as C

    import torch

    from mwdiem_amd import _lib, scene
    from mwdiem_amd.stepper import Stepper
    sc, v, sp = _pour()
    with pytest.raises(_lib.MwdError):  # one scene only
        Stepper([sc, sc], spawn=sp)
    with pytest.raises(_lib.MwdError):  # at least the walls are active
        Stepper(sc, spawn=scene.Spawn(n_active=-1, every=10, per=5))
    S = Stepper(sc, spawn=sp)
    with pytest.raises(_lib.MwdError):  # no sorted cell pass with a spawn schedule
        S.set_tuning(cell_pass=_lib.CELLS_SORTED)
    S.set_law_seeds()
    lt = {k: torch.zeros_like(S.t[k]) for k in _lib.TAN_ARRAYS}
    ptrs = (C.c_void_p * len(lt))(*[lt[k].data_ptr() for k in _lib.TAN_ARRAYS])
    with pytest.raises(_lib.MwdError):  # nor the adjoint step
        S.step_adjoint(ptrs, torch.zeros(3, dtype=torch.float64, device=S.device))

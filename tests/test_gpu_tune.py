"""The autotuner: it restores the state it times on, leaves a valid choice
behind, and a widened cell (physical law) finds the same contacts."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import random_motion

pytestmark = pytest.mark.gpu


def test_autotune_restores_state_and_changes_no_result(gpu, dense_phys):
    from mwdiem_amd import _lib
    from mwdiem_amd.stepper import Stepper
    from mwdiem_amd.tune import SPACE, autotune
    sc = dense_phys
    st = random_motion(O.initial_state(sc), sc, 9, vscale=2000.0, wscale=200.0)
    S, R = Stepper(sc), Stepper(sc)
    S.set_state(st)
    R.set_state(st)
    before = S.numpy_state()
    res = autotune(S, steps=3)
    after = S.numpy_state()
    for k in before:
        assert np.array_equal(before[k], after[k]), k
    t = S.tuning()
    assert t == res["tuning"]
    for k, vals in SPACE.items():
        assert t[k] in vals, (k, t[k])
    assert len(res["trials"]) >= 2
    S.step(15)
    R.step(15)
    S.check()
    R.check()
    a, b = S.numpy_state(), R.numpy_state()
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    assert R.tuning()["cell_pass"] == _lib.CELLS_TABLE


def _contact_set(S, st):
    S.set_state(st)
    S.candidates()
    S.fine_contact()
    S.check()
    c = S.contacts()
    pairs = S.pairs()[c["pair"]]
    return {tuple(sorted(map(int, p))) for p in pairs}


def test_wider_cells_find_the_same_contacts(gpu, dense_phys, impact_scene):
    from mwdiem_amd.stepper import Stepper
    from mwdiem_amd.tune import regrid, tune_cell_scale
    sc = dense_phys
    for seed in (None, 3):
        st = O.initial_state(sc)
        if seed is not None:
            st = random_motion(st, sc, seed)
        base = _contact_set(Stepper(sc), st)
        assert base
        for f in (1.3, 2.0):
            wide = regrid(sc, f)
            assert wide.ncx < sc.ncx
            assert _contact_set(Stepper(wide), st) == base, (seed, f)
    best, ms = tune_cell_scale([dense_phys, impact_scene], scales=(1.0, 1.5), steps=3)
    assert best in (1.0, 1.5) and set(ms) == {1.0, 1.5} and all(v > 0 for v in ms.values())

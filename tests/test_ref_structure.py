"""The reference-structure CPU path (oracle/ref_structure.py: the reference's
host loop in pure Python, the CPU baseline) against the C++ oracle: the same
arithmetic in the same order, so the same state bit for bit after steps with
contacts, on both laws."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from oracle import ref_structure as R


@pytest.mark.parametrize("mode", [0, 1])
def test_reference_structure_equals_oracle(mode):
    from mwdiem_amd import scene
    base = scene.make_scene(30, (6.0, 6.0, 12.0), 17, law=scene.Law4(mode=mode))
    sc = scene.compact_dense(base, gap=-0.55)
    a = O.initial_state(sc)
    rng = np.random.Generator(np.random.PCG64(3))
    W = sc.n_walls
    a["v"][W:] = a["v12"][W:] = rng.normal(size=(sc.n_particles, 3)) * 100.0
    b = {k: v.copy() for k, v in a.items()}
    n_o = n_r = 0
    for _ in range(3):
        n_o += int((O.step(a, sc)[1] == 1).sum())
        n_r += R.step(b, sc)
    assert n_o == n_r and n_o > 10
    for k in O.STATE_KEYS:
        assert np.array_equal(a[k], b[k]), k

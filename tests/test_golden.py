"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py):
the oracle must keep reproducing them (CPU), and the GPU path must reproduce
them (GPU)."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

GOLD = Path(__file__).resolve().parent / "golden"
NAMES = ["dense_ref", "dense_phys"]


def _load(name):
    from mwdiem_amd.scene import SceneData
    d = dict(np.load(GOLD / f"{name}.npz"))
    sc = SceneData.from_arrays(d)
    st = {k[3:]: d[k].copy() for k in d if k.startswith("in_")}
    return sc, st, d


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    sc, st, d = _load(name)
    pairs = O.candidates(st, sc)
    assert np.array_equal(pairs, d["pairs"])
    coll, depth, nrm, ca, cb, F, T, est = O.contact_state(pairs, st, sc)
    assert np.array_equal(coll, d["collision"])
    for a, k in ((depth, "depth"), (nrm, "normal"), (ca, "cpA"), (cb, "cpB"), (F, "force"), (T, "torque")):
        assert np.array_equal(a, d[k]), k
    for _ in range(int(d["steps"])):
        O.step(st, sc)
    for k in ("com", "v", "w", "q", "q12", "qall", "verts"):
        assert np.array_equal(st[k], d[f"out_{k}"]), k


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_reproduces_golden(gpu, name):
    from mwdiem_amd.stepper import Stepper
    from tests.helpers import max_rel
    sc, st, d = _load(name)
    S = Stepper(sc)
    S.set_state(st)
    S.candidates()
    S.fine_contact()
    S.check()
    assert np.array_equal(S.pairs(), d["pairs"])
    c = S.contacts()
    assert np.array_equal(c["collision"], d["collision"])
    idx = np.nonzero(d["collision"] == 1)[0]
    for k in ("depth", "normal", "cpA", "cpB"):
        assert max_rel(c[k], d[k][idx]) <= 1e-9, k
    S2 = Stepper(sc)
    S2.set_state(st)
    S2.step(int(d["steps"]))
    S2.check()
    got = S2.numpy_state()
    W = sc.n_walls
    for k in ("com", "q", "qall"):
        assert max_rel(got[k][W:], d[f"out_{k}"][W:]) <= 1e-10, k

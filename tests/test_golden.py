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


def test_oracle_flags_unresolved_epa():
    """A pair met while settling the pile (tests/golden/make_epa_unresolved.py)
    whose EPA ends on a deleted face: the oracle returns the placeholder
    distance 3e20 and status bit 4."""
    d = np.load(GOLD / "epa_unresolved.npz")
    c, depth, n, ca, cb, st = O.gjk_epa(d["VA"], d["VB"], d["comA"], d["comB"])
    assert c == 1 and depth == 3e20 and st & 4
    assert c == int(d["collision"]) and depth == float(d["depth"]) and st == int(d["status"])


def test_oracle_physical_variant_resolves_that_pair():
    """The physical law's GJK/EPA variant (faces turned away from an interior
    point) resolves the same pair: depth and normal of the Minkowski
    difference's nearest facet (scipy Qhull), status 0."""
    from scipy.spatial import ConvexHull
    d = np.load(GOLD / "epa_unresolved.npz")
    A, B = d["VA"], d["VB"]
    eq = ConvexHull((A[:, None, :] - B[None, :, :]).reshape(-1, 3)).equations
    i = int(np.argmax(eq[:, 3]))
    c, depth, n, ca, cb, st = O.gjk_epa(A, B, d["comA"], d["comB"], robust=True)
    assert c == 1 and st == 0
    assert abs(depth + eq[i, 3]) <= 1e-12 and np.abs(n - eq[i, :3]).max() <= 1e-6


@pytest.mark.gpu
def test_gpu_physical_law_resolves_the_unresolved_pair(gpu):
    """The same pair in a two-hull scene on the physical law: the GPU's
    variant resolves it as the oracle's does (no MWD_F_EPA_UNRESOLVED) and one
    GPU step equals the oracle's."""
    from mwdiem_amd import _lib, scene
    from mwdiem_amd.stepper import Stepper
    from tests.helpers import max_rel
    d = np.load(GOLD / "epa_unresolved.npz")
    sc = scene.make_scene(2, (40.0, 40.0, 60.0), 5, npts=(8, 20), size=(0.6, 0.9),
                          law=scene.Law4(mode=scene.PHYSICAL), name="epa-unresolved")
    W = sc.n_walls
    vmax = max(sc.vmax, d["VA"].shape[0], d["VB"].shape[0])
    if vmax > sc.vmax:
        sc.verts = np.concatenate([sc.verts, np.zeros((sc.n_rows, vmax - sc.vmax, 3))], axis=1)
        sc.vmax = vmax
    for r, V, c in ((W, d["VA"], d["comA"]), (W + 1, d["VB"], d["comB"])):
        sc.verts[r] = 0.0
        sc.verts[r, :V.shape[0]] = V
        sc.nverts[r] = V.shape[0]
        sc.com[r] = c
    st = O.initial_state(sc)
    S = Stepper(sc)
    S.set_state(st)
    ref = {k: v.copy() for k, v in st.items()}
    pairs, coll = O.step(ref, sc)
    S.step(1)
    s = S.check()
    assert not s["flags"] & _lib.F_EPA_UNRESOLVED
    assert np.count_nonzero(coll == 1) >= 1
    got = S.numpy_state()
    for k in ("com", "v", "w", "q"):
        assert max_rel(got[k][W:], ref[k][W:]) <= 1e-12, k
    assert np.abs(got["v"][W:]).max() < 1e3  # no blow-up

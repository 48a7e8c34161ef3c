"""Shallow contacts on which the reference's GJK/EPA fails, and the physical
law's variant that resolves them (tests/golden/flat_contacts.npz, written by
tests/golden/make_flat_contacts.py from a poured pile of 4,096 hulls).

On these pairs GJK ends on a tetrahedron with the origin within 1e-6 mm of
the planes CorrectFlat tests (gpu_fine_contact_detection.cu:256-268), the
fourth vertex is re-supported along S1 x S2 (.cu:1119-1128), the origin ends
just outside the tetrahedron, AddTriangle turns that face's normal inwards
(.cu:285-290: it orients by the sign of the origin's distance) and EPA
returns a face about a hull's size deep: a penalty force 10^4 - 10^6 times
the true one. The physical law's GJK/EPA (flat tetrahedra judged by their
own volume, faces turned away from an interior point, an absolute stopping
tolerance) returns the Minkowski difference's nearest facet.

CPU: both variants of the oracle against scipy's Qhull of the Minkowski
difference. GPU: the HIP path on the fixture's scene (walls plus the pairs,
physical law) against the oracle, contact by contact and over a full step.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest
from scipy.spatial import ConvexHull

from oracle import oracle as O
from tests.helpers import max_rel

FIX = Path(__file__).resolve().parent / "golden" / "flat_contacts.npz"


def _load():
    from mwdiem_amd.scene import SceneData
    from mwdiem_amd.stepper import STATE_KEYS
    d = dict(np.load(FIX))
    sc = SceneData.from_arrays(d)
    st = {k: d["state_" + k].copy() for k in STATE_KEYS}
    return sc, st, d["pairs"]


def _bodies(sc, st, a, b):
    return st["verts"][a, :sc.nverts[a]], st["verts"][b, :sc.nverts[b]]


def _nearest_facet(A, B):
    """Depth and outward normal of the Minkowski difference A - B's facet
    nearest the origin (the penetration depth and direction)."""
    M = (A[:, None, :] - B[None, :, :]).reshape(-1, 3)
    eq = ConvexHull(M).equations
    assert np.all(eq[:, 3] < 0)  # the origin is inside: the hulls overlap
    i = int(np.argmax(eq[:, 3]))
    return float(-eq[i, 3]), eq[i, :3]


def test_reference_epa_fails_on_the_fixture():
    """The quirk as the reference has it: CorrectFlat fires and the depth is
    more than 10^4 times the nearest facet's."""
    sc, st, pairs = _load()
    assert sc.law_p.mode == 1 and sc.law_w.mode == 1
    c, it, flat, es, ei = O.gjk_trace(pairs, st["com"], st["verts"], sc.nverts)
    assert np.all(c == 1) and np.all(flat == 1) and np.all(es == 0)
    for a, b in pairs:
        A, B = _bodies(sc, st, a, b)
        true, _ = _nearest_facet(A, B)
        _, depth, *_ = O.gjk_epa(A, B, st["com"][a], st["com"][b])
        assert depth > 1e4 * true and depth > 0.5


def test_physical_epa_finds_the_nearest_facet():
    """The physical law's variant: the nearest facet's depth (to 1e-9 mm) and
    normal, the contact points on the two support planes along it, and
    cpA - cpB = depth * n."""
    sc, st, pairs = _load()
    c, it, flat, es, ei = O.gjk_trace(pairs, st["com"], st["verts"], sc.nverts, robust=True)
    assert np.all(c == 1) and np.all(es == 0)
    for a, b in pairs:
        A, B = _bodies(sc, st, a, b)
        true, nt = _nearest_facet(A, B)
        coll, depth, n, cpA, cpB, est = O.gjk_epa(A, B, st["com"][a], st["com"][b], robust=True)
        assert coll == 1 and est == 0
        assert abs(depth - true) <= 1e-9 and 0 < depth < 1e-5
        assert np.abs(n - nt).max() <= 1e-6
        # A's deepest point along n and B's along -n
        assert abs(n @ cpA - (A @ n).max()) <= 1e-9
        assert abs(n @ cpB - (B @ n).min()) <= 1e-9
        assert np.abs((cpA - cpB) - depth * n).max() <= 1e-9


def test_physical_step_stays_bounded():
    """One step of the fixture's scene: with the physical variant no hull is
    kicked (every hull's speed change stays below 1 mm/s, gravity's is
    9.8e-3), where the reference's depths would add 10^3 mm/s and more."""
    sc, st, pairs = _load()
    ref = {k: v.copy() for k, v in st.items()}
    O.step(ref, sc)
    W = sc.n_walls
    dv = np.linalg.norm(ref["v"][W:] - st["v"][W:], axis=1)
    assert dv.max() < 1.0
    # the same pairs with the reference's EPA: the force of the failed depth
    Kn = sc.law_p.Kn
    for a, b in pairs:
        A, B = _bodies(sc, st, a, b)
        _, depth, *_ = O.gjk_epa(A, B, st["com"][a], st["com"][b])
        assert Kn * depth * sc.dt / sc.mass[a] > 1e3


@pytest.mark.gpu
@pytest.mark.parametrize("fused", [0, 1])
def test_gpu_flat_contacts_match_oracle(gpu, fused):
    """The HIP path's physical-law GJK/EPA on the fixture: the same contacts,
    depths, normals and points as the oracle (bitwise), and one full step
    within 1e-12 of the oracle's."""
    from mwdiem_amd.stepper import Stepper
    sc, st, pairs = _load()
    S = Stepper(sc)
    S.set_tuning(epa_fused=fused)
    S.set_state(st)
    S.candidates()
    S.fine_contact()
    S.check()
    got = S.contacts()
    coll, depth, nrm, ca, cb, F, T, est = O.contact_state(S.pairs(), st, sc)
    idx = np.nonzero(coll == 1)[0]
    assert np.array_equal(got["collision"], coll)
    assert np.array_equal(got["pair"], idx)
    pl = S.pairs()[idx]
    for a, b in pairs:  # every fixture pair is among the contacts
        assert np.any((pl[:, 0] == a) & (pl[:, 1] == b)) or np.any((pl[:, 0] == b) & (pl[:, 1] == a))
    for name, ref in (("depth", depth), ("normal", nrm), ("cpA", ca), ("cpB", cb)):
        assert np.array_equal(got[name], ref[idx]), name
    assert np.array_equal(got["epa_status"], est[idx])
    assert depth[idx].max() < 1e-5
    S.set_state(st)
    S.step(1)
    S.check()
    ref = {k: v.copy() for k, v in st.items()}
    O.step(ref, sc)
    out = S.numpy_state()
    W = sc.n_walls
    for k in ("com", "v", "w", "q"):
        assert max_rel(out[k][W:], ref[k][W:], scale=1.0) <= 1e-12, k
    S.close()

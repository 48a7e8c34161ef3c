"""The reference's rare GJK/EPA branches, on hull pairs that reach them
(tests/golden/rare_paths.npz, written by tests/golden/make_rare_paths.py):

  * CorrectFlat (gpu_fine_contact_detection.cu:256-268, 1119-1128): the
    fourth simplex vertex is re-supported along S1 x S2 when both simplex
    determinants fall under 1e-6 -- hulls about 1e-3 mm across;
  * EPA's expansion cap (.cu:443-447): 25 expansions, status bit 1;
  * GJK's 50-iteration tetrahedron cap (.cu:1104,1287, Q13): collision stays
    -1 -- the origin lies on a support plane of the Minkowski difference.

CPU: the oracle reproduces the fixture and each case is pinned by what the
geometry fixes (scipy Qhull of the Minkowski difference). GPU: the HIP path
through mwd_contact_pairs gives the same collision codes and EPA status, and
depth, normal and contact points within 1e-9 relative.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest
from scipy.spatial import ConvexHull

from oracle import oracle as O

FIX = Path(__file__).resolve().parent / "golden" / "rare_paths.npz"


def _fix():
    return dict(np.load(FIX))


def _bodies(d, i):
    a, b = d["pairs"][i]
    return d["verts"][a, :d["nverts"][a]], d["verts"][b, :d["nverts"][b]]


def test_oracle_reproduces_rare_paths():
    d = _fix()
    c, it, fl, es, ei = O.gjk_trace(d["pairs"], d["com"], d["verts"], d["nverts"])
    for k, v in (("collision", c), ("gjk_iters", it), ("flat", fl), ("epa_status", es), ("epa_iters", ei)):
        assert np.array_equal(v, d[k]), k
    fam = d["family"]
    assert np.all(fl[fam == "flat"] == 1)
    assert np.all((es[fam == "epa_cap"] & 1) == 1) and np.all(ei[fam == "epa_cap"] == 26)
    assert np.all(c[fam == "gjk_cap"] == -1) and np.all(it[fam == "gjk_cap"] == 50)


def test_correct_flat_cases_against_minkowski_hull():
    """After CorrectFlat, EPA still finds the penetration of the Minkowski
    difference: depth within EPA's 1e-5 of the nearest facet, contact points
    on both support planes."""
    d = _fix()
    for i in np.flatnonzero(d["family"] == "flat"):
        A, B = _bodies(d, i)
        M = (A[:, None, :] - B[None, :, :]).reshape(-1, 3)
        off = ConvexHull(M).equations[:, 3]
        assert np.all(off < 0)  # overlapping
        true = float(np.min(-off))
        c, depth, n, cpA, cpB, st = O.gjk_epa(A, B, A.mean(0), B.mean(0))
        assert c == 1 and st == 0
        assert true / (1 + 1.1e-5) <= depth <= true * (1 + 1e-12), i
        s = np.abs(A).max()
        assert np.allclose(cpA - cpB, depth * n, rtol=0, atol=1e-12 * s)
        assert abs(cpA @ n - (A @ n).max()) <= 1e-12 * s and abs(cpB @ n - (B @ n).min()) <= 1e-12 * s


def test_epa_cap_cases_are_inner_bounds():
    """Stopped after 25 expansions, EPA's polytope lies inside the Minkowski
    difference, so its closest face is no farther than the true depth."""
    d = _fix()
    for i in np.flatnonzero(d["family"] == "epa_cap"):
        A, B = _bodies(d, i)
        M = (A[:, None, :] - B[None, :, :]).reshape(-1, 3)
        true = float(np.min(-ConvexHull(M).equations[:, 3]))
        c, depth, n, cpA, cpB, st = O.gjk_epa(A, B, A.mean(0), B.mean(0))
        assert c == 1 and st & 1
        assert 0.0 < depth <= true * (1 + 1e-12), i
        assert abs(np.linalg.norm(n) - 1) < 1e-12


def test_gjk_cap_cases_touch_a_support_plane():
    """The -1 code comes from pairs whose Minkowski difference has the origin
    on a support plane (the strict dot(d, S) < 0 exit never fires): apart by
    Qhull, yet max over A-B along the placement direction is 0 to rounding."""
    d = _fix()
    for i in np.flatnonzero(d["family"] == "gjk_cap"):
        A, B = _bodies(d, i)
        M = (A[:, None, :] - B[None, :, :]).reshape(-1, 3)
        assert np.max(ConvexHull(M).equations[:, 3]) > 0  # the origin is outside
        u = d["touch_dir"][i]
        assert abs(np.max(M @ u)) <= 1e-9 * np.abs(M).max()


@pytest.mark.gpu
def test_gpu_rare_paths_match_oracle(gpu):
    from tests.helpers import gpu_contact_pairs, max_rel
    d = _fix()
    for fam in ("flat", "epa_cap", "gjk_cap"):
        print(f"rare paths: {fam}", flush=True)
    g = gpu_contact_pairs(d["verts"], d["nverts"], d["com"], d["pairs"])
    assert np.array_equal(g["collision"], d["collision"])
    assert np.array_equal(g["epa_status"], d["epa_status"])
    hit = d["collision"] == 1
    for i in np.flatnonzero(hit):
        for k in ("depth", "normal", "cpA", "cpB"):
            assert max_rel(g[k][i], d[k][i], scale=1e-300 + np.abs(d[k][i]).max()) <= 1e-9, (d["family"][i], k)
    # no force from an undecided pair (Q13)
    cap = np.flatnonzero(d["family"] == "gjk_cap")
    assert np.all(g["force"][d["pairs"][cap].ravel()] == 0.0)


@pytest.mark.gpu
def test_gpu_out_of_box_reset(gpu, small_scene):
    """gpu_Verlet.cu:338-352 (Q15): a particle whose centre leaves the box in
    a step has its velocities zeroed and its translation undone, keeps its
    rotation, and dx_all keeps the step's dx -- the drop-in Verlet_GPU
    against the oracle, with a few particles thrown out through each face."""
    from mwdiem_amd.dropin import gpu_Verlet
    from tests.helpers import max_rel, random_motion
    sc = small_scene
    W, N = sc.n_walls, sc.n_rows
    st = random_motion(O.initial_state(sc), sc, 5)
    box = sc.box
    out_rows = []
    for k, ax in enumerate((0, 0, 1, 1, 2, 2)):
        r = W + 3 * k
        hi = k % 2 == 0
        st["com"][r, ax] = box[2 * ax] - 1e-3 if hi else box[2 * ax + 1] + 1e-3
        st["v12"][r, :] = 0.0
        st["v12"][r, ax] = 5e3 if hi else -5e3   # 5e-3 mm per 1e-6 s step: out through the face
        out_rows.append(r)
    F = np.zeros((N, 3))
    T = np.zeros((N, 3))
    ref = {k: v.copy() for k, v in st.items()}
    O.verlet(ref, F, T, sc, W, N)
    got = {k: v.copy() for k, v in st.items()}
    gpu_Verlet.Verlet_GPU(got["v"], got["v12"], got["w12bf"], got["q12"], got["q"], got["qall"], got["wbf"],
                          got["w"], F, T, got["verts"], got["com"], sc.dt, sc.mass.reshape(-1, 1), N, sc.Iinv,
                          sc.I, got["dxall"], sc.vmax, *sc.box, W, N)
    for r in out_rows:
        for k in ("v", "v12", "w", "wbf", "w12bf"):
            assert np.array_equal(got[k][r], [0.0, 0.0, 0.0]), (r, k)
        assert np.allclose(got["com"][r], st["com"][r], rtol=0, atol=1e-12)
        assert np.any(got["q"][r] != st["q"][r])          # rotation kept
        assert np.any(got["dxall"][r] != 0.0)             # dx_all keeps dx
    for k in ("com", "v", "v12", "w", "q", "q12", "qall", "dxall", "verts"):
        assert max_rel(got[k][W:], ref[k][W:]) <= 1e-12, k

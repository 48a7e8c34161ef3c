"""Parity of the HIP path (through the C ABI) with the oracle, on the GPU.

Bars (from the project's parity contract):
  * candidate pair lists and collision codes: identical (order included);
  * penetration depth, normal, contact points: within 1e-9 relative;
  * per-step integrator outputs: within 1e-12 relative (the only differences
    allowed are the last-ulp differences of sin/cos between the device math
    library and glibc);
  * positions and quaternions after 100 full steps: within 1e-8 relative.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import max_rel, random_motion

pytestmark = pytest.mark.gpu


def _stepper(sc, **kw):
    from mwdiem_amd.stepper import Stepper
    return Stepper(sc, **kw)


def test_native_library_is_loaded_once(gpu):
    with open("/proc/self/maps") as f:
        maps = f.read()
    assert "libmwdiem_hip.so" in maps
    hip = {ln.split()[-1] for ln in maps.splitlines() if "libamdhip64" in ln}
    assert len(hip) == 1, f"more than one HIP runtime mapped: {hip}"


@pytest.mark.parametrize("seed", [1, 2])
def test_verlet_dropin_one_step(gpu, small_scene, seed):
    from mwdiem_amd.dropin import gpu_Verlet
    sc = small_scene
    W, N = sc.n_walls, sc.n_rows
    st = random_motion(O.initial_state(sc), sc, seed)
    rng = np.random.Generator(np.random.PCG64(100 + seed))
    F = rng.normal(size=(N, 3)) * 1e-3
    T = rng.normal(size=(N, 3)) * 1e-6
    ref = {k: v.copy() for k, v in st.items()}
    O.verlet(ref, F, T, sc, W, N)
    got = {k: v.copy() for k, v in st.items()}
    out = gpu_Verlet.Verlet_GPU(got["v"], got["v12"], got["w12bf"], got["q12"], got["q"], got["qall"], got["wbf"],
                                got["w"], F, T, got["verts"], got["com"], sc.dt, sc.mass.reshape(-1, 1), N, sc.Iinv,
                                sc.I, got["dxall"], sc.vmax, *sc.box, W, N)
    # reference return order, gpu_Verlet.py:89-92, and in-place update
    order = ("v12", "q", "wbf", "w12bf", "w", "q12", "v", "com", "verts", "qall", "dxall")
    for name, arr in zip(order, out):
        assert arr is got[name]
    for k in order:
        err = max_rel(got[k][W:], ref[k][W:])
        assert err <= 1e-12, (k, err)
    # walls untouched
    for k in order:
        assert np.array_equal(got[k][:W], st[k][:W]), k


def test_verlet_100_steps(gpu, small_scene):
    from mwdiem_amd.dropin import gpu_Verlet
    sc = small_scene
    W, N = sc.n_walls, sc.n_rows
    st = random_motion(O.initial_state(sc), sc, 7)
    rng = np.random.Generator(np.random.PCG64(7))
    F = rng.normal(size=(N, 3)) * 1e-3
    T = rng.normal(size=(N, 3)) * 1e-6
    ref = {k: v.copy() for k, v in st.items()}
    got = {k: v.copy() for k, v in st.items()}
    for _ in range(100):
        O.verlet(ref, F, T, sc, W, N)
        gpu_Verlet.Verlet_GPU(got["v"], got["v12"], got["w12bf"], got["q12"], got["q"], got["qall"], got["wbf"],
                              got["w"], F, T, got["verts"], got["com"], sc.dt, sc.mass.reshape(-1, 1), N, sc.Iinv,
                              sc.I, got["dxall"], sc.vmax, *sc.box, W, N)
    for k in ("com", "q", "qall", "v", "w"):
        err = max_rel(got[k][W:], ref[k][W:])
        assert err <= 1e-8, (k, err)


@pytest.mark.parametrize("which", ["small", "dense", "dense_phys", "impact"])
def test_candidates_identical(gpu, small_scene, dense_scene, dense_phys, impact_scene, which):
    sc = {"small": small_scene, "dense": dense_scene, "dense_phys": dense_phys, "impact": impact_scene}[which]
    S = _stepper(sc)
    S.candidates()
    s = S.check()
    got = S.pairs()
    ref = O.candidates(O.initial_state(sc), sc)
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)
    nw = O.wall_pairs(sc.com, sc.n_walls, sc.n_rows, sc.wall_n, sc.wall_c, sc.wall_thr).shape[0]
    assert s["n_wall_pairs"] == nw


def _crowded(sc, factor=2.0):
    """The scene on a grid of cells `factor` times wider, so cells hold many
    members (the set-order emulation beyond four members, sorting of the
    members claimed in any order)."""
    from dataclasses import replace
    from mwdiem_amd import scene
    cell, ncx, ncy, ncz, kmin, kspan = scene.grid_for(sc.box, factor * sc.max_particlesize)
    return replace(sc.copy(), cell=cell, ncx=ncx, ncy=ncy, ncz=ncz, kmin=kmin, kspan=kspan)


@pytest.mark.parametrize("which", ["dense", "dense_phys", "crowded", "crowded_phys"])
def test_cell_passes_agree(gpu, dense_scene, dense_phys, which):
    """Cells from the key tables (the default) and from the radix sort of the
    keys give the same candidate list, order included, step after step while
    particles change cells; the first list also matches the oracle."""
    from mwdiem_amd import _lib
    sc = {"dense": dense_scene, "dense_phys": dense_phys, "crowded": _crowded(dense_scene),
          "crowded_phys": _crowded(dense_phys)}[which]
    st = random_motion(O.initial_state(sc), sc, 21, vscale=3000.0, wscale=300.0)
    A, B = _stepper(sc), _stepper(sc)
    assert A.tuning()["cell_pass"] == _lib.CELLS_TABLE
    B.set_tuning(cell_pass=_lib.CELLS_SORTED)
    assert B.tuning()["cell_pass"] == _lib.CELLS_SORTED
    A.set_state(st)
    B.set_state(st)
    A.candidates()
    assert np.array_equal(A.pairs(), O.candidates(st, sc))
    if which.startswith("crowded"):
        assert A.status()["n_cells"] * 4 < sc.n_particles  # many members per cell
    # the same workspaces on shifted copies of the bed: every particle changes
    # cell, so the key tables are reused with other keys occupied
    W = sc.n_walls
    rng = np.random.Generator(np.random.PCG64(7))
    for i in range(4):
        d = sc.cell * rng.uniform(-0.5, 0.5, size=3)
        si = {k: v.copy() for k, v in st.items()}
        si["com"][W:] += d
        si["verts"][W:] += d
        A.set_state(si)
        B.set_state(si)
        A.candidates()
        B.candidates()
        sa, sb = A.check(), B.check()
        assert sa["n_cells"] == sb["n_cells"], i
        assert np.array_equal(A.pairs(), B.pairs()), i
        if i == 0:
            assert np.array_equal(A.pairs(), O.candidates(si, sc))
    for i in range(8):
        A.step(1)
        B.step(1)
        A.candidates()
        B.candidates()
        sa, sb = A.check(), B.check()
        assert sa["n_cells"] == sb["n_cells"], i
        assert np.array_equal(A.pairs(), B.pairs()), i
    a, b = A.numpy_state(), B.numpy_state()
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_large_scene_cells_and_candidates(gpu, large_scene):
    """At full size (100,000 hulls, 890k candidates): the candidate list of
    the key-table cells equals the oracle's, order included, and the sorted
    cells give the same list and the same state after 10 steps."""
    from mwdiem_amd import _lib
    sc = large_scene
    st = O.initial_state(sc)
    A = _stepper(sc)
    A.candidates()
    A.check()
    ref = O.candidates(st, sc)
    got = A.pairs()
    assert got.shape == ref.shape and np.array_equal(got, ref)
    B = _stepper(sc)
    B.set_tuning(cell_pass=_lib.CELLS_SORTED)
    B.candidates()
    B.check()
    assert np.array_equal(B.pairs(), ref)
    A.step(10)
    B.step(10)
    A.check()
    B.check()
    assert np.array_equal(A.pairs(), B.pairs())
    a, b = A.numpy_state(), B.numpy_state()
    for k in ("com", "q", "v", "w", "verts"):
        assert np.array_equal(a[k], b[k]), k


def test_tuning_choices_change_nothing(gpu, dense_phys):
    """Every launch choice of the tuning interface gives the same state, bit
    for bit; out-of-range choices are refused."""
    from mwdiem_amd import _lib
    sc = dense_phys
    st = random_motion(O.initial_state(sc), sc, 5, vscale=2000.0, wscale=200.0)
    ref = _stepper(sc)
    ref.set_state(st)
    ref.step(20)
    ref.check()
    want = ref.numpy_state()
    assert ref.status()["n_contacts"] > 0
    variants = [dict(gjk_group=4), dict(gjk_group=16), dict(gjk_blocks=64), dict(gjk_blocks=8192),
                dict(sphere_block=64), dict(sphere_block=1024), dict(verlet_block=64), dict(verlet_block=128),
                dict(cell_items=0), dict(cell_items=2), dict(cell_items=4),
                dict(cell_pass=_lib.CELLS_SORTED, gjk_group=16), dict(verlet_waves=3), dict(verlet_waves=4),
                dict(vertex_pack=1), dict(vertex_pack=1, verlet_block=64, verlet_waves=4), dict(epa_blocks=16),
                dict(epa_blocks=128), dict(epa_fused=0), dict(epa_fused=0, gjk_group=16),
                dict(gjk_group=16, gjk_blocks=64), dict(fill_cells=16), dict(fill_cells=4),
                dict(fill_cells=4, cell_pass=_lib.CELLS_SORTED), dict(gjk_buckets=1),
                dict(gjk_buckets=1, gjk_group=4), dict(gjk_buckets=1, epa_fused=0), dict(fill_split=4),
                dict(fill_split=16, fill_cells=4), dict(fill_split=4, cell_pass=_lib.CELLS_SORTED)]
    for v in variants:
        S = _stepper(sc)
        S.set_tuning(**v)
        S.set_state(st)
        S.step(20)
        S.check()
        got = S.numpy_state()
        for k in want:
            assert np.array_equal(got[k], want[k]), (v, k)
        S.close()
    S = _stepper(sc)
    for bad in (dict(gjk_group=5), dict(gjk_blocks=8), dict(sphere_block=100), dict(cell_pass=7),
                dict(verlet_block=96), dict(cell_items=3), dict(verlet_waves=2), dict(vertex_pack=2),
                dict(epa_blocks=8), dict(epa_fused=2), dict(fill_cells=8), dict(gjk_buckets=2),
                dict(fill_split=2)):
        with pytest.raises(_lib.MwdError):
            S.set_tuning(**bad)


@pytest.mark.parametrize("group,fused", [(4, 1), (4, 0), (8, 1), (8, 0)])
def test_gjk_classes_change_nothing(gpu, group, fused):
    """GJK per class of vertex count on hulls of 6 to 24 vertices (classes
    of 4 or 8 vertices per lane step, fused with EPA or not): the same
    state as one launch, bit for bit, after contact-bearing steps."""
    from mwdiem_amd import scene
    base = scene.make_scene(80, (12.0, 12.0, 30.0), 21, npts=(6, 24), law=scene.Law4(mode=scene.PHYSICAL),
                            name="mixed-80")
    sc = scene.compact_dense(base, gap=-0.55)
    assert len(set(sc.nverts[sc.n_walls:].tolist())) > 8
    st = random_motion(O.initial_state(sc), sc, 3, vscale=2000.0, wscale=200.0)
    out = []
    for buckets in (0, 1):
        S = _stepper(sc)
        S.set_tuning(gjk_group=group, epa_fused=fused, gjk_buckets=buckets)
        S.set_state(st)
        S.step(15)
        stt = S.check()
        assert stt["n_contacts"] > 0
        out.append(S.numpy_state())
        S.close()
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k]), k


@pytest.mark.parametrize("pack", [0, 1])
def test_graph_replay_equals_eager(gpu, dense_phys, pack):
    """Steps replayed from a captured graph give the eager steps' state bit
    for bit, replay after replay (with eager work and host reads between the
    replays); the graph holds the wall scan on the side stream and its
    join."""
    sc = dense_phys
    st = random_motion(O.initial_state(sc), sc, 7, vscale=2000.0, wscale=200.0)
    A = _stepper(sc)
    B = _stepper(sc)
    for S in (A, B):
        S.set_tuning(vertex_pack=pack)
        S.set_state(st)
    B.capture(5)
    seen = 0
    for r in range(4):
        A.step(5)
        B.run(5)
        A.check()
        B.check()
        seen += A.status()["n_contacts"]
        a, b = A.numpy_state(), B.numpy_state()
        for k in a:
            assert np.array_equal(a[k], b[k]), (r, k)
    assert seen > 0


@pytest.mark.parametrize("seed,phys", [(3, False), (4, False), (3, True)])
def test_fine_contact_matches_oracle(gpu, dense_scene, dense_phys, seed, phys):
    sc = dense_phys if phys else dense_scene
    S = _stepper(sc)
    st = random_motion(O.initial_state(sc), sc, seed)
    S.set_state(st)
    S.candidates()
    S.fine_contact()
    S.check()
    pairs = S.pairs()
    c = S.contacts()
    coll, depth, nrm, ca, cb, F, T, est = O.contact_state(pairs, st, sc)
    assert np.array_equal(c["collision"], coll)
    idx = np.nonzero(coll == 1)[0]
    assert idx.size > 20
    assert np.array_equal(c["pair"], idx)
    assert max_rel(c["depth"], depth[idx]) <= 1e-9
    for name, ref in (("normal", nrm), ("cpA", ca), ("cpB", cb)):
        assert max_rel(c[name], ref[idx]) <= 1e-9, name
    assert np.array_equal(c["epa_status"], est[idx])


def test_fine_contact_is_bitwise(gpu, dense_scene):
    """Same operator order, no contraction, correctly rounded division and
    square root on both sides: the EPA results are bit-identical."""
    sc = dense_scene
    S = _stepper(sc)
    st = random_motion(O.initial_state(sc), sc, 13)
    S.set_state(st)
    S.candidates()
    S.fine_contact()
    S.check()
    c = S.contacts()
    coll, depth, nrm, ca, cb, F, T, est = O.contact_state(S.pairs(), st, sc)
    idx = np.nonzero(coll == 1)[0]
    for name, ref in (("depth", depth), ("normal", nrm), ("cpA", ca), ("cpB", cb)):
        assert np.array_equal(c[name], ref[idx]), name


def test_step_reduction_matches_oracle(gpu, dense_scene):
    """One full step: the fused reduction + integrator sees the same forces."""
    sc = dense_scene
    S = _stepper(sc)
    st = random_motion(O.initial_state(sc), sc, 5)
    S.set_state(st)
    ref = {k: v.copy() for k, v in st.items()}
    O.step(ref, sc)
    S.step(1)
    S.check()
    got = S.numpy_state()
    W = sc.n_walls
    for k in ("com", "v", "v12", "w", "wbf", "w12bf", "q", "q12", "qall", "dxall"):
        assert max_rel(got[k][W:], ref[k][W:]) <= 1e-12, k


def test_reduction_with_more_contacts_than_a_bucket(gpu):
    """A particle with more contacts than its bucket holds (its extra records
    go through the overflow list): the step still sums in ascending pair
    order and matches the oracle."""
    from mwdiem_amd import _lib, scene
    base = scene.make_scene(300, (20.0, 20.0, 40.0), 13, law=scene.Law4(mode=scene.PHYSICAL), name="crowd")
    base.inner = (10.0, 10.0, 40.0)  # packed into a column: many neighbours in 3-D
    sc = scene.compact_dense(base, gap=-0.65)
    W = sc.n_walls
    st = O.initial_state(sc)
    pairs = O.candidates(st, sc)
    coll = O.contact_state(pairs, st, sc)[0]
    cnt = np.bincount(pairs[coll == 1].ravel(), minlength=sc.n_rows)
    r = W + int(np.argmax(cnt[W:]))
    n = sc.nverts[r]
    sc.verts[r, :n] = sc.com[r] + 2.0 * (sc.verts[r, :n] - sc.com[r])  # a large hull among small ones
    st = O.initial_state(sc)
    pairs = O.candidates(st, sc)
    coll = O.contact_state(pairs, st, sc)[0]
    cnt = np.bincount(pairs[coll == 1].ravel(), minlength=sc.n_rows)
    assert cnt[r] > _lib.ROW_CONTACTS
    S = _stepper(sc, pair_cap=8 * len(pairs))
    S.set_state(st)
    ref = {k: v.copy() for k, v in st.items()}
    O.step(ref, sc)
    S.step(1)
    S.check()
    got = S.numpy_state()
    for k in ("com", "v", "v12", "w", "wbf", "w12bf", "q", "q12", "qall", "dxall"):
        assert max_rel(got[k][W:], ref[k][W:]) <= 1e-12, k


def test_timing_mode_changes_nothing(gpu, dense_phys):
    """With timing on, the single-kernel stages are launched with start/stop
    events (hipExtLaunchKernelGGL): the results stay bit-identical and the
    kernel times are positive. Steps with EPA inside the GJK launch add no
    EPA time, also when they alternate with steps that launch it."""
    A = _stepper(dense_phys)
    B = _stepper(dense_phys)
    B.set_timing(True)
    A.step(30)
    B.step(15)
    B.set_tuning(epa_fused=0)
    B.step(10)
    B.set_tuning(epa_fused=1)
    B.step(5)
    kt = B.kernel_times()
    st = B.stage_times()
    B.step(4)
    kt_fused = B.kernel_times()
    B.stage_times()
    B.set_timing(False)
    A.step(4)
    A.check()
    B.check()
    a, b = A.numpy_state(), B.numpy_state()
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    assert kt["steps"] == 30 and st["steps"] == 30
    # the sphere test runs inside the emission when bounding radii are kept
    assert kt["ms"]["k_sphere"] == 0.0
    assert all(v > 0 for k, v in kt["ms"].items() if k != "k_sphere"), kt
    assert kt_fused["steps"] == 4 and kt_fused["ms"]["k_epa_force"] == 0.0 and kt_fused["ms"]["k_gjk"] > 0, kt_fused


def _run_both(sc, st, steps, check_pairs=5):
    S = _stepper(sc)
    S.set_state(st)
    n_contacts = 0
    for i in range(steps):
        pairs_ref, coll_ref = O.step(st, sc)
        S.step(1)
        n_contacts += int((coll_ref == 1).sum())
        if i < check_pairs:
            assert np.array_equal(S.pairs(), pairs_ref), i
    S.check()
    return S.numpy_state(), n_contacts


def test_100_steps_reference_law(gpu, small_scene):
    """Reference law, lattice bed with random motion: keys, candidates and the
    integrator over 100 steps."""
    sc = small_scene
    st = random_motion(O.initial_state(sc), sc, 11, vscale=300.0, wscale=30.0, quats=False)
    got, _ = _run_both(sc, st, 100)
    W = sc.n_walls
    for k in ("com", "q", "qall", "v", "w"):
        assert max_rel(got[k][W:], st[k][W:]) <= 1e-8, k
    nv = sc.nverts
    for r in range(W, sc.n_rows):
        assert max_rel(got["verts"][r, :nv[r]], st["verts"][r, :nv[r]], scale=1.0) <= 1e-8


def test_100_steps_physical_law_with_impacts(gpu, impact_scene):
    """Physical law, deposited hulls thrown at each other: contacts every step."""
    sc = impact_scene
    W = sc.n_walls
    st = O.initial_state(sc)
    rng = np.random.Generator(np.random.PCG64(5))
    st["v"][W:] = rng.normal(size=(sc.n_particles, 3)) * 1000.0
    st["v12"][W:] = st["v"][W:]
    got, n_contacts = _run_both(sc, st, 100)
    assert n_contacts > 100
    for k in ("com", "q", "qall"):
        assert max_rel(got[k][W:], st[k][W:]) <= 1e-8, k


def test_contact_dropin_matches_oracle(gpu, dense_scene):
    """Contact_detection_GPU with the driver's FP32 per-pair arrays."""
    from mwdiem_amd.dropin import gpu_fine_contact_detection as FC
    sc = dense_scene
    st = random_motion(O.initial_state(sc), sc, 9)
    pairs = O.candidates(st, sc)
    P, vmax, N, W = pairs.shape[0], sc.vmax, sc.n_rows, sc.n_walls
    a, b = pairs[:, 0], pairs[:, 1]
    f32 = np.float32
    vertsA = st["verts"][a].astype(f32)
    vertsB = st["verts"][b].astype(f32)
    for k in range(P):  # zero padding as the driver gathers (MWDiem.py:850-859)
        vertsA[k, sc.nverts[a[k]]:] = 0
        vertsB[k, sc.nverts[b[k]]:] = 0
    args = dict(comA=st["com"][a].astype(f32), comB=st["com"][b].astype(f32), vA=st["v"][a].astype(f32),
                vB=st["v"][b].astype(f32), wA=st["w"][a].astype(f32), wB=st["w"][b].astype(f32),
                nA=sc.nverts[a].astype(np.int32), nB=sc.nverts[b].astype(np.int32), idA=a.astype(np.int32),
                idB=b.astype(np.int32), disp=np.zeros((P, 4), f32), qA=st["q"][a].astype(f32),
                qB=st["q"][b].astype(f32), isA=sc.isize[a].astype(f32), isB=sc.isize[b].astype(f32))
    force = np.zeros((N, 3), f32)
    torque = np.zeros((N, 3), f32)
    L, Lw = sc.law_p, sc.law_w
    out = FC.Contact_detection_GPU(args["comA"], args["comB"], vertsA, vertsB, args["vA"], args["vB"], args["wA"],
                                   args["wB"], args["nA"], args["nB"], args["idA"], args["idB"], args["disp"], L.Kn,
                                   L.en, L.et, L.mu, Lw.Kn, Lw.en, Lw.et, Lw.mu, W, force, torque, vmax, args["qA"],
                                   args["qB"], sc.dt, args["isA"], args["isB"])
    fo, to, disp, coll = out
    assert fo is force and to is torque and disp is args["disp"]
    Fr = np.zeros((N, 3))
    Tr = np.zeros((N, 3))
    ref = O.contact_pairs(args["comA"], args["comB"], vertsA, vertsB, args["vA"], args["vB"], args["wA"], args["wB"],
                          args["nA"], args["nB"], args["idA"], args["idB"], np.zeros((P, 4)), args["qA"], args["qB"],
                          args["isA"], args["isB"], [L.Kn, L.en, L.et, L.mu], [Lw.Kn, Lw.en, Lw.et, Lw.mu], W,
                          sc.dt, Fr, Tr)
    assert np.array_equal(coll, ref["collision"])
    assert (coll == 1).sum() > 20
    assert np.array_equal(force, Fr.astype(f32))
    assert np.array_equal(torque, Tr.astype(f32))
    assert np.array_equal(disp, ref["disp"].astype(f32))


def test_hash_dropin_matches_reference_order(gpu, dense_scene):
    from mwdiem_amd.dropin import LIBRARY_Coarse_contact_detection as CC
    sc = dense_scene
    keys = O.cell_keys(sc.com, sc.cell, sc.ncx, sc.ncy)
    start = [[1, 2]]
    out = CC.Linked_cell_hash(None, sc.ncx, sc.ncy, sc.n_walls, sc.n_rows, start, keys.reshape(-1, 1))
    assert out is start
    ref = O.linked_cell_hash_py(keys, sc.n_walls, sc.n_rows, sc.ncx, sc.ncy)
    assert out[1:] == ref


def test_batched_equals_separate(gpu, small_scene, dense_scene):
    """Several scenes in one Stepper: each scene's pairs (in order) and state
    are bit-identical to a Stepper of that scene alone."""
    from mwdiem_amd import scene
    from mwdiem_amd.stepper import Stepper
    sc3 = scene.compact_dense(scene.make_scene(40, (8.0, 8.0, 20.0), 12, law=scene.Law4(Kn=3e4, en=0.3, et=0.4,
                                                                                        mu=0.5)), gap=-0.5)
    scs = [dense_scene, small_scene, sc3]
    B = Stepper(scs)
    singles = [Stepper(s) for s in scs]
    for _ in range(3):
        B.step(1)
        for s in singles:
            s.step(1)
    B.check()
    pb = B.pairs()
    for i, s in enumerate(singles):
        s.check()
        rows = B.scene_rows(i)
        sel = (pb[:, 0] >= rows.start) & (pb[:, 0] < rows.stop)
        assert np.array_equal(pb[sel] - rows.start, s.pairs()), i
        gb = B.numpy_state()
        gs = s.numpy_state()
        for k in ("com", "q", "v", "w", "verts"):
            assert np.array_equal(gb[k][rows], gs[k]), (i, k)


@pytest.mark.parametrize("fused", [1, 0])
def test_persistent_tangential_history_matches_oracle(gpu, dense_phys, fused):
    """Persistent tangential history (the history the driver means to keep,
    keyed min+max^2): 60 steps of the dense bed with random motion against
    the oracle carrying the same history, within 1e-8, with EPA inside the
    GJK launch and as its own; the history changes the result (the
    reference's reset mode gives another state)."""
    from mwdiem_amd.stepper import Stepper
    sc = dense_phys
    st0 = random_motion(O.initial_state(sc), sc, 8, vscale=800.0, wscale=100.0)
    st = {k: v.copy() for k, v in st0.items()}
    hist = {}
    S = Stepper(sc, persistent_history=True)
    R = Stepper(sc)
    S.set_tuning(epa_fused=fused)
    S.set_state(st0)
    R.set_state(st0)
    n_kept = 0
    for i in range(60):
        O.step(st, sc, history=hist)
        n_kept += len(hist)
        S.step(1)
        R.step(1)
    S.check()
    R.check()
    got, reset = S.numpy_state(), R.numpy_state()
    W = sc.n_walls
    assert n_kept > 100
    for k in ("com", "q", "v", "w"):
        assert max_rel(got[k][W:], st[k][W:]) <= 1e-8, k
    assert np.abs(got["v"][W:] - reset["v"][W:]).max() > 1e-6 * np.abs(reset["v"][W:]).max()


def test_sphere_survivors_in_crowded_cells(gpu):
    """Crowded cells give blocks of the pair emission thousands of rounds, so
    the survivors of the sphere test go through the emission's buffered
    appends many times: their count equals the wall pairs the wall's slab
    does not separate from the particle's bounding sphere plus the cell pairs
    whose bounding spheres meet, recounted on the host from the emitted
    candidates, the radii, the centres and the walls' vertices. A second pass
    over the same candidates (the standalone sphere-test kernel, walls tested
    too) counts the pairs whose spheres meet among all of them."""
    from dataclasses import replace
    from mwdiem_amd import scene
    base = scene.make_scene(3000, (60.0, 60.0, 120.0), 5, name="crowd-3000")
    sc = scene.compact_dense(base, gap=-0.55)
    cell, ncx, ncy, ncz, kmin, kspan = scene.grid_for(sc.box, 2.0 * sc.max_particlesize)
    sc = replace(sc.copy(), cell=cell, ncx=ncx, ncy=ncy, ncz=ncz, kmin=kmin, kspan=kspan)
    st = random_motion(O.initial_state(sc), sc, 3, vscale=500.0, wscale=100.0)
    S = _stepper(sc, pair_cap=4_000_000, contact_cap=200_000)
    S.set_state(st)
    S.candidates()
    S.fine_contact()
    stat = S.check()
    pairs = S.pairs()
    W = stat["n_wall_pairs"]
    # cells of ~20 members: 64-cell blocks of the emission hold hundreds of
    # 256-pair rounds each
    assert (len(pairs) - W) / stat["n_cells"] > 1024
    com = S.numpy_state()["com"]
    rad = S.t["radius"].cpu().numpy()

    def meet(pp):
        a, b = pp[:, 0], pp[:, 1]
        d = com[b] - com[a]
        n2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        rr = rad[a] + rad[b]
        return ~(n2 > rr * rr)

    cell_meet = int(meet(pairs[W:]).sum())
    assert cell_meet > 1000
    # wall pairs: the particle's sphere and the wall's slab |n.(x - c)| <= h
    # apart by more than 1e-6 (1 + |dd| + h) are decided in the wall scan
    verts = S.numpy_state()["verts"]
    wall_keep = 0
    for j, k in pairs[:W]:
        n, c = sc.wall_n[k], sc.wall_c[k]
        dd = abs(float(n @ (com[j] - c)))
        h = float(np.abs((verts[k, :sc.nverts[k]] - c) @ n).max())
        wall_keep += not (dd - h - rad[j] > 1e-6 * (1.0 + dd + h))
    assert 0 < wall_keep < W
    assert stat["n_sphere_pass"] == wall_keep + cell_meet
    S.fine_contact()  # a second pass: k_sphere over the stored candidates
    assert S.check()["n_sphere_pass"] == int(meet(pairs).sum())


@pytest.mark.parametrize("fused", [1, 0])
def test_contact_overflow_is_reported(gpu, dense_phys, fused):
    """More contacts than contact_cap: the contacts past the cap are dropped
    without a write out of bounds and the step reports MWD_E_CAPACITY, with
    EPA inside the GJK launch and as its own."""
    from mwdiem_amd import _lib
    sc = dense_phys
    st = random_motion(O.initial_state(sc), sc, 5, vscale=2000.0, wscale=200.0)
    ref = _stepper(sc)
    ref.set_state(st)
    ref.step(20)
    need = ref.status()["n_contacts"]
    assert need > 2
    S = _stepper(sc, contact_cap=2)
    S.set_tuning(epa_fused=fused)
    S.set_state(st)
    with pytest.raises(_lib.MwdError, match="contact_cap"):
        for _ in range(20):
            S.step(1)
            S.check()

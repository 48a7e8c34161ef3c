"""Pins of the oracle (CPU): each check is a closed form, a brute force, an
invariant or the interpreter's own behaviour -- never the oracle's formula
retyped -- chosen so that a plausible slip (dropped term, wrong sign or index,
swapped operand) fails one of them."""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np
import pytest
from scipy.spatial import ConvexHull

from oracle import oracle as O


# ------------------------------------------------------------ quaternions
def test_quaternion_kat_reference_test_py():
    # test.py:16-31 prints (1,2,3,4)x(5,6,7,8); the Hamilton product is (-60,12,30,24)
    assert np.array_equal(O.quat_mul([1, 2, 3, 4], [5, 6, 7, 8]), [-60, 12, 30, 24])


def test_quaternion_kat_verlet_dead_code():
    # gpu_Verlet.cu:406-426 multiplies (1,2,-3,-1)x(-1,3,43,-1); by hand:
    # w = -1 - (6 - 129 + 1) = 121, v = (3,43,-1) - (2,-3,-1) + (46,-1,95)
    assert np.array_equal(O.quat_mul([1, 2, -3, -1], [-1, 3, 43, -1]), [121, 47, 45, 95])


def _rotmat(q):
    w, x, y, z = np.asarray(q, float) / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def test_quaternion_triple_product_is_scaled_rotation():
    # q v q^-1 with the conjugate as inverse = |q|^2 R(q) v (gpu_Verlet.cu:419-426 uses a=(1,2,-3,-1))
    a = np.array([1.0, 2.0, -3.0, -1.0])
    ainv = a * np.array([1, -1, -1, -1])
    v = np.array([1.0, 4.0, -2.0])
    got = O.quat_tripl(a, v, ainv)
    assert np.allclose(got, np.dot(a, a) * _rotmat(a) @ v, rtol=0, atol=1e-12)


# ------------------------------------------------------------ Verlet
class _Sc:
    """Minimal scene for the integrator: one wall row, particles after it."""

    def __init__(self, n, dt=1e-6, I=(1.0, 2.0, 3.0), mass=2.0, vmax=4, box=1e9, mode=0):
        self.n_walls = 1
        self.n_rows = n + 1
        self.vmax = vmax
        self.dt = dt
        self.mass = np.full(self.n_rows, mass)
        self.I = np.tile(np.diag(I).astype(float), (self.n_rows, 1, 1))
        self.Iinv = np.tile(np.diag(1.0 / np.asarray(I, float)), (self.n_rows, 1, 1))
        self.box = np.array([box, -box, box, -box, box, -box], float)
        self.nverts = np.full(self.n_rows, vmax, np.int32)
        self.com = np.zeros((self.n_rows, 3))
        self.verts = np.zeros((self.n_rows, vmax, 3))

        class L:
            pass
        self.law_p = L()
        self.law_p.mode = mode


def _state(sc):
    N = sc.n_rows
    st = {k: np.zeros((N, 3)) for k in ("v", "v12", "w", "wbf", "w12bf", "dxall")}
    for k in ("q", "q12", "qall"):
        st[k] = np.zeros((N, 4))
        st[k][:, 0] = 1.0
    st["com"] = sc.com.copy()
    st["verts"] = sc.verts.copy()
    return st


def test_verlet_free_fall_closed_form():
    sc = _Sc(3, dt=1e-4)
    rng = np.random.default_rng(0)
    sc.com[1:] = rng.normal(size=(3, 3))
    st = _state(sc)
    g = np.array([0.0, 0.0, -9810.0])
    F = np.tile(sc.mass[:, None] * g, (1, 1))
    T = np.zeros_like(F)
    k = 57
    for _ in range(k):
        O.verlet(st, F, T, sc, 1, sc.n_rows)
    dt = sc.dt
    # v_{1/2} after k steps: k g dt; x_k = x_0 + g dt^2 k(k+1)/2; velocity = v12 + g dt
    assert np.allclose(st["v12"][1:], k * g * dt, rtol=1e-12, atol=0)
    assert np.allclose(st["com"][1:], sc.com[1:] + g * dt * dt * k * (k + 1) / 2, rtol=1e-12, atol=1e-15)
    assert np.allclose(st["v"][1:], (k + 1) * g * dt, rtol=1e-12, atol=0)
    assert np.allclose(st["dxall"][1:], g * dt * dt * k * (k + 1) / 2, rtol=1e-12, atol=1e-18)
    assert np.array_equal(st["q"][1:], np.tile([1.0, 0, 0, 0], (3, 1)))


@pytest.mark.parametrize("mode", [0, 1])
def test_verlet_torque_free_spin_closed_form(mode):
    Om, dt = 40.0, 1e-4
    sc = _Sc(1, dt=dt, mode=mode)
    sc.verts[1] = np.array([[1.0, 0, 0], [0, 2.0, 0], [0, 0, 1.0], [-1.0, -1.0, 0.5]])
    st = _state(sc)
    for k in ("w", "wbf", "w12bf"):
        st[k][1] = [0.0, 0.0, Om]
    F = np.zeros((2, 3))
    k = 25
    for _ in range(k):
        O.verlet(st, F, F, sc, 1, 2, rot_mode=mode)
    # alpha = 0 about a principal axis: q12 = rot(k Om dt), q = rot(k Om dt + Om dt / 2)
    a12 = k * Om * dt
    aq = a12 + Om * dt / 2
    assert np.allclose(st["q12"][1], [math.cos(a12 / 2), 0, 0, math.sin(a12 / 2)], rtol=0, atol=1e-14)
    assert np.allclose(st["q"][1], [math.cos(aq / 2), 0, 0, math.sin(aq / 2)], rtol=0, atol=1e-14)
    assert np.allclose(st["w"][1], [0, 0, Om], rtol=1e-14)
    # vertices: the reference rotates by the absolute q every step (the angles
    # add up, gpu_Verlet.cu:279-305); the physical mode by the increment only
    total = sum(j * Om * dt + Om * dt / 2 for j in range(1, k + 1)) if mode == 0 else aq
    c, s = math.cos(total), math.sin(total)
    Rz = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    assert np.allclose(st["verts"][1], sc.verts[1] @ Rz.T, rtol=0, atol=1e-12)


def test_verlet_unit_quaternion_and_zero_rotation_invariants():
    sc = _Sc(20, dt=1e-5, I=(0.7, 1.1, 1.9))
    st = _state(sc)
    rng = np.random.default_rng(3)
    st["w"][1:11] = rng.normal(size=(10, 3)) * 30
    st["wbf"][1:11] = rng.normal(size=(10, 3)) * 30
    st["w12bf"][1:11] = st["wbf"][1:11]
    F = np.zeros((21, 3))
    T = np.zeros((21, 3))
    T[1:11] = rng.normal(size=(10, 3))
    for _ in range(200):
        O.verlet(st, F, T, sc, 1, 21)
    assert np.max(np.abs(np.linalg.norm(st["q"][1:11], axis=1) - 1)) < 1e-13
    # zero torque and zero spin keep q = (1,0,0,0) exactly
    assert np.array_equal(st["q"][11:], np.tile([1.0, 0, 0, 0], (10, 1)))


def test_verlet_out_of_box_reset():
    # gpu_Verlet.cu:338-352: velocities zeroed, translation undone, rotation kept, dx_all keeps dx
    sc = _Sc(1, dt=1e-3, box=1.0)
    sc.com[1] = [0.99, 0.0, 0.0]
    sc.verts[1] = sc.com[1] + np.eye(4, 3)
    st = _state(sc)
    st["v12"][1] = [100.0, 0, 0]
    st["w"][1] = [0, 0, 5.0]
    st["wbf"][1] = [0, 0, 5.0]
    st["w12bf"][1] = [0, 0, 5.0]
    F = np.zeros((2, 3))
    O.verlet(st, F, F, sc, 1, 2)
    # (x + dx) + (-dx): back to x up to one rounding
    assert np.allclose(st["com"][1], sc.com[1], rtol=0, atol=1e-15)
    for k in ("v", "v12", "w", "wbf", "w12bf"):
        assert np.array_equal(st[k][1], [0, 0, 0]), k
    assert np.allclose(st["dxall"][1], [0.1, 0, 0])
    assert st["q"][1][3] != 0.0


# ------------------------------------------------------------ GJK / EPA
def _box(lo, hi):
    return np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])], float)


@pytest.mark.parametrize("delta", [0.05, 0.2, 0.45])
def test_overlapping_boxes_closed_form(delta):
    A = _box((0, 0, 0), (1, 1, 1))
    B = _box((1 - delta, 0.13, 0.21), (2.3 - delta, 1.17, 1.29))
    c, depth, n, cpA, cpB, st = O.gjk_epa(A, B, A.mean(0), B.mean(0))
    assert c == 1 and st == 0
    assert abs(depth - delta) <= 1e-5 * delta
    assert abs(abs(n[0]) - 1) < 1e-9 and abs(n[1]) < 1e-9 and abs(n[2]) < 1e-9
    # contact points lie on the faces that touch: A's x=1 face, B's x=1-delta face
    assert abs(cpA[0] - 1) < 1e-9 and abs(cpB[0] - (1 - delta)) < 1e-9


def test_separated_boxes():
    A = _box((0, 0, 0), (1, 1, 1))
    for shift in ((1.5, 0, 0), (0, 1.01, 0.2), (-0.2, 0.3, -1.2)):
        B = A + np.array(shift)
        c, *_ = O.gjk_epa(A, B, A.mean(0), B.mean(0))
        assert c == 0, shift


def _hull_pair(rng, sep):
    def hull(n, scale):
        p = rng.normal(size=(n, 3))
        p /= np.linalg.norm(p, axis=1, keepdims=True)
        p *= scale
        return p[ConvexHull(p).vertices]
    A = hull(int(rng.integers(6, 20)), rng.uniform(0.6, 1.2, size=3))
    B = hull(int(rng.integers(6, 20)), rng.uniform(0.6, 1.2, size=3))
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    return A, B + d * sep


def test_gjk_epa_against_minkowski_hull():
    """Brute force: the Minkowski difference A-B via Qhull. Overlap iff the
    origin is inside; the penetration depth is the smallest facet distance to
    the origin. EPA stops at 1e-5 relative (.cu:439), so its depth lies in
    [true/(1+1e-5), true]."""
    rng = np.random.default_rng(7)
    n_hit = 0
    for trial in range(120):
        A, B = _hull_pair(rng, rng.uniform(0.2, 2.6))
        M = (A[:, None, :] - B[None, :, :]).reshape(-1, 3)
        h = ConvexHull(M)
        off = h.equations[:, 3]  # n.x + off <= 0 inside
        inside = bool(np.all(off < 0))
        c, depth, n, cpA, cpB, st = O.gjk_epa(A, B, A.mean(0), B.mean(0))
        if np.min(np.abs(off)) < 1e-9:
            continue  # origin on the boundary: either answer is right
        assert c in (0, 1)
        assert (c == 1) == inside, trial
        if c == 1 and st == 0:
            n_hit += 1
            true = float(np.min(-off))
            assert true / (1 + 1.1e-5) <= depth <= true * (1 + 1e-12), (trial, depth, true)
            assert abs(np.linalg.norm(n) - 1) < 1e-12
            # the contact points realise the penetration: cpA - cpB is the
            # Minkowski point depth*n; cpA lies on A's support plane along n,
            # cpB on B's support plane along -n, each inside its own hull
            assert np.allclose(cpA - cpB, depth * n, rtol=0, atol=1e-12)
            assert abs(cpA @ n - (A @ n).max()) <= 1e-12 and abs(cpB @ n - (B @ n).min()) <= 1e-12
            for P_, H in ((cpA, A), (cpB, B)):
                eq = ConvexHull(H).equations
                assert np.all(eq[:, :3] @ P_ + eq[:, 3] <= 1e-12), trial
    assert n_hit >= 20


def test_gjk_epa_swap_symmetry():
    rng = np.random.default_rng(11)
    for _ in range(40):
        A, B = _hull_pair(rng, rng.uniform(0.1, 1.2))
        c1, d1, n1, *_ = O.gjk_epa(A, B, A.mean(0), B.mean(0))
        c2, d2, n2, *_ = O.gjk_epa(B, A, B.mean(0), A.mean(0))
        assert c1 == c2
        if c1 == 1:
            assert abs(d1 - d2) <= 2.2e-5 * max(d1, d2)


def test_flop_counter_on_an_early_exit():
    # separated along d1: one support call then exit. Flops: d1 (3), support
    # dots 5 per vertex of A and B, S (3), dot(d1,S) (5); negation is free.
    A = _box((0, 0, 0), (1, 1, 1))
    B = A + np.array([3.0, 0.0, 0.0])
    assert O.pair_flops(A, B, A.mean(0), B.mean(0)) == 3 + 5 * 8 + 5 * 8 + 3 + 5


# ------------------------------------------------------------ contact law
def _collide(mode, en, mass, steps, v0=100.0, Kn=1e4, mu=0.3, dt=1e-7):
    """Two boxes (or a box and a wall) approaching head-on along x, rotation
    frozen; returns the relative normal velocity ratio out/in."""
    from mwdiem_amd import scene
    W = 1
    sc = _Sc(2, dt=dt, I=(1e30, 1e30, 1e30), mass=mass, vmax=8, mode=mode)
    sc.n_walls = W
    sc.com = np.array([[0, 0, -50.0], [0, 0, 0], [1.0, 0.13, 0.21]])
    sc.verts = np.stack([_box((-10, -10, -60), (10, 10, -40)), _box((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5)),
                         _box((0.5, -0.37, -0.29), (1.5, 0.63, 0.71))])
    sc.com[1:] = sc.verts[1:].mean(axis=1)
    sc.isize = np.full(3, 1e30)
    sc.law_p = scene.Law4(Kn=Kn, en=en, et=0.5, mu=mu, mode=mode)
    sc.law_w = sc.law_p
    sc.wall_n = np.zeros((3, 3))
    sc.wall_c = np.zeros((3, 3))
    sc.wall_thr = -1.0
    sc.fext = np.zeros((3, 3))
    sc.text = np.zeros((3, 3))
    sc.cell = np.array([100.0, 100.0, 100.0])
    sc.ncx = sc.ncy = 1
    st = _state(sc)
    st["v"][1] = st["v12"][1] = [v0 / 2, 0, 0]
    st["v"][2] = st["v12"][2] = [-v0 / 2, 0, 0]
    pairs = np.array([[1, 2]], np.int32)
    for _ in range(steps):
        coll, *_r, F, T, _e = O.contact_state(pairs, st, sc)
        O.verlet(st, F, T, sc, 1, 3, nrot=sc.nverts)
    return (st["v"][2, 0] - st["v"][1, 0]) / v0


def test_physical_law_restitution():
    # linear spring-dashpot: the normal restitution coefficient is en
    for en in (0.3, 0.7):
        e = _collide(1, en, mass=2e-6, steps=4000)
        assert abs(e - en) < 0.02 * en + 1e-3, (en, e)


def test_reference_law_restitution_with_mu_as_mass():
    # the reference's damping uses mu where the effective mass belongs (.cu:690-691):
    # with particle masses 2*mu (effective mass mu) the restitution is en
    mu = 0.3
    # contact lasts pi/sqrt(Kn/mu) = 5.4e-4 s = 540 steps of 1e-6 s
    e = _collide(0, 0.5, mass=2 * mu, steps=1500, v0=100.0, Kn=1e7, mu=mu, dt=1e-6)
    assert abs(e - 0.5) < 0.02, e


# ------------------------------------------------------------ coarse stage
def test_cell_keys_truncate_toward_zero():
    rng = np.random.default_rng(1)
    com = rng.uniform(-30, 60, size=(500, 3))
    cell = np.array([2.3, 1.7, 3.1])
    keys = O.cell_keys(com, cell, 17, 9)
    xc = (com / cell).astype(int)  # MWDiem.py:813-815
    assert np.array_equal(keys, xc[:, 0] + xc[:, 1] * 17 + xc[:, 2] * 17 * 9)
    assert O.cell_keys(np.array([[-0.5, 0, 0]]), np.ones(3), 4, 4)[0] == 0


def test_wall_pairs_brute_force():
    rng = np.random.default_rng(2)
    W, N = 4, 300
    com = rng.uniform(-5, 15, size=(N, 3))
    wn = rng.normal(size=(N, 3))
    wn /= np.linalg.norm(wn, axis=1, keepdims=True)
    wc = rng.uniform(0, 10, size=(N, 3))
    got = O.wall_pairs(com, W, N, wn, wc, 3.0)
    ref = [[j, k] for k in range(W) for j in range(W, N) if abs(np.dot(wn[k], com[j] - wc[k])) < 3.0]
    assert got.tolist() == ref


def test_set_order_is_cpythons():
    rng = np.random.default_rng(5)
    for _ in range(3000):
        n = int(rng.integers(1, 80))
        ids = np.sort(rng.choice(int(rng.choice([80, 600, 50000])), size=n, replace=False))
        s = set()
        for x in ids.tolist():
            s.add(x)
        assert O.set_order(ids) == list(s)


@pytest.mark.parametrize("ncx,ncy", [(7, 5), (1, 4), (3, 1), (1, 1)])
def test_linked_cell_hash_equals_python_restatement(ncx, ncy):
    rng = np.random.default_rng(ncx * 10 + ncy)
    N, W = 400, 5
    keys = rng.integers(-20, 120, size=N)
    got = O.linked_cell_hash(keys, W, N, ncx, ncy).tolist()
    assert got == O.linked_cell_hash_py(keys, W, N, ncx, ncy)


def test_linked_cell_hash_brute_force_multiset():
    """Order-free pin: pair (a,b) appears once per offset o with
    key_b - key_a == o (A in the lower cell), and once for same-cell pairs."""
    from collections import Counter
    rng = np.random.default_rng(9)
    N, W, ncx, ncy = 250, 3, 6, 4
    keys = rng.integers(0, 6 * 4 * 5, size=N)
    offs = [1, ncx, ncx * ncy, ncx * ncy + 1, ncx * ncy + 1 + ncx, ncx * ncy + ncx, ncx + 1]
    ref = Counter()
    for a in range(W, N):
        for b in range(W, N):
            if a == b:
                continue
            if keys[a] == keys[b] and a < b:
                ref[frozenset((a, b))] += 1
            for o in offs:
                if keys[b] - keys[a] == o:
                    ref[(a, b)] += 1
    got = Counter()
    for a, b in O.linked_cell_hash(keys, W, N, ncx, ncy).tolist():
        if keys[a] == keys[b]:
            got[frozenset((a, b))] += 1
        else:
            got[(a, b)] += 1
    assert got == ref


def test_full_stencil_is_the_half_shell():
    from collections import Counter
    rng = np.random.default_rng(4)
    ncx, ncy = 5, 6
    N = 200
    xyz = rng.integers(1, 4, size=(N, 3))
    keys = xyz[:, 0] + xyz[:, 1] * ncx + xyz[:, 2] * ncx * ncy
    got = Counter(map(frozenset, O.linked_cell_hash(keys, 0, N, ncx, ncy, stencil=1).tolist()))
    ref = Counter()
    for a in range(N):
        for b in range(a + 1, N):
            if np.all(np.abs(xyz[a] - xyz[b]) <= 1):
                ref[frozenset((a, b))] += 1
    assert got == ref


# ------------------------------------------------------------ tangential history
def test_tangential_history_stick_then_slip():
    """Two overlapping boxes held in place, B sliding past A along y at v;
    the history fed back each step (the persistent mode), physical law.
    While sticking the stored displacement grows as k*v*dt and the
    tangential force stays under mu*|Fn|; from the first slip on the
    tangential force equals mu*|Fn| and opposes the sliding, and the stored
    displacement stays within one step's travel of the slip value (at the
    Coulomb limit a step may count as sticking by rounding and advance it by
    v*dt, the next one resets it) (.cu:770-811 with qA = 1; the slip
    displacement with the sign of the physical law)."""
    from mwdiem_amd import scene
    dt, v, mu = 1e-6, 1000.0, 0.3
    sc = _Sc(2, dt=dt, I=(1e30, 1e30, 1e30), mass=2e-6, vmax=8, mode=1)
    sc.com = np.zeros((3, 3))
    sc.verts = np.stack([_box((-10, -10, -60), (10, 10, -40)), _box((0, 0, 0), (1, 1, 1)),
                         _box((0.95, 0.13, 0.21), (1.95, 1.13, 1.21))])
    sc.com = sc.verts.mean(axis=1)
    sc.isize = np.full(3, 1e30)
    sc.law_p = scene.Law4(Kn=1e4, en=0.5, et=0.5, mu=mu, mode=1)
    sc.law_w = sc.law_p
    st = _state(sc)
    st["v"][2] = [0.0, -v, 0.0]   # relative velocity of A against B: +v along y
    pairs = np.array([[1, 2]], np.int32)
    disp = np.zeros((1, 4))
    slip_at = None
    prev = None
    for k in range(1, 41):
        coll, _, _, _, _, F, _, _, new = O.contact_state(pairs, st, sc, dispt=disp, with_disp=True)
        assert coll[0] == 1
        Fn, Ft = abs(F[1, 0]), np.hypot(F[1, 1], F[1, 2])
        if slip_at is None and Ft < mu * Fn * (1 - 1e-12):
            assert abs(new[0, 1] - k * v * dt) <= 1e-12 * k * v * dt
            assert abs(new[0, 0]) < 1e-15 and abs(new[0, 2]) < 1e-15
        else:
            slip_at = slip_at or k
            assert abs(Ft - mu * Fn) <= 1e-12 * Fn
            assert F[1, 1] < 0.0  # on A, against A's sliding along +y
            if prev is None:
                prev = new[0, :3].copy()  # the slip displacement
            assert np.all(np.abs(new[0, :3] - prev) <= v * dt * (1 + 1e-9))
        disp = new
    assert slip_at is not None and 3 < slip_at < 40
    # without history (the reference as it runs) the same step gives a smaller tangential force
    _, _, _, _, _, F0, _, _ = O.contact_state(pairs, st, sc)
    assert np.hypot(F0[1, 1], F0[1, 2]) < np.hypot(F[1, 1], F[1, 2])


# ------------------------------------------------------------ contact points, torques, tangential law
def _tip_into_face(delta, tip=(0.3, -0.2)):
    """A tetrahedron standing on its tip, pressed delta deep into the top
    face (z = 0) of a wide box. The minimum translation is along z: EPA's
    contact point on A is the tip, on B the tip's projection onto the face."""
    tx, ty = tip
    A = np.array([[tx, ty, -delta],
                  [tx + 1.0, ty, 1.0 - delta],
                  [tx - 0.5, ty + 0.866, 1.0 - delta],
                  [tx - 0.5, ty - 0.866, 1.0 - delta]])
    B = _box((-2.0, -2.0, -1.0), (2.0, 2.0, 0.0))
    return A, B


def _edge_on_face(delta, x0=-0.4, x1=0.7, y=0.15):
    """A wedge whose lowest edge (x0..x1 at height -delta, y fixed) is pressed
    into the same box face."""
    A = np.array([[x0, y, -delta], [x1, y, -delta],
                  [x0, y - 0.5, 0.8 - delta], [x1, y - 0.5, 0.8 - delta],
                  [x0, y + 0.5, 0.8 - delta], [x1, y + 0.5, 0.8 - delta]])
    B = _box((-2.0, -2.0, -1.0), (2.0, 2.0, 0.0))
    return A, B


@pytest.mark.parametrize("delta", [0.02, 0.1, 0.3])
def test_contact_points_tip_into_face(delta):
    A, B = _tip_into_face(delta)
    c, depth, n, cpA, cpB, st = O.gjk_epa(A, B, A.mean(0), B.mean(0))
    assert c == 1 and st == 0
    assert abs(depth - delta) <= 1e-5 * delta
    assert np.allclose(n, [0, 0, -1], atol=1e-12)
    assert np.allclose(cpA, A[0], rtol=0, atol=1e-12)                 # the tip
    assert np.allclose(cpB, [A[0, 0], A[0, 1], 0.0], rtol=0, atol=1e-12)  # its projection on the face


@pytest.mark.parametrize("delta", [0.05, 0.2])
def test_contact_points_edge_on_face(delta):
    x0, x1, y = -0.4, 0.7, 0.15
    A, B = _edge_on_face(delta, x0, x1, y)
    c, depth, n, cpA, cpB, st = O.gjk_epa(A, B, A.mean(0), B.mean(0))
    assert c == 1 and st == 0
    assert abs(depth - delta) <= 1e-5 * delta
    assert np.allclose(n, [0, 0, -1], atol=1e-12)
    # on the edge segment, and straight above it on the face
    assert abs(cpA[1] - y) < 1e-12 and abs(cpA[2] + delta) < 1e-12 and x0 - 1e-12 <= cpA[0] <= x1 + 1e-12
    assert np.allclose(cpB, [cpA[0], cpA[1], 0.0], rtol=0, atol=1e-12)


def _pair_scene(A, B, comA, comB, law, isA=1.0, isB=1.0, mA=1.0, mB=1.0):
    """Two bodies as rows 1 and 2 of a scene (row 0: an unused wall) for
    O.contact_state; state at rest, q = 1."""
    from mwdiem_amd import scene
    vmax = max(len(A), len(B))
    sc = _Sc(2, I=(1.0, 1.0, 1.0), mass=1.0, vmax=vmax, mode=law.mode)
    sc.verts = np.zeros((3, vmax, 3))
    sc.verts[1, :len(A)] = A
    sc.verts[2, :len(B)] = B
    sc.nverts = np.array([vmax, len(A), len(B)], np.int32)
    sc.com = np.array([[0.0, 0.0, 100.0], comA, comB], float)
    sc.isize = np.array([1.0, isA, isB])
    sc.mass = np.array([1.0, mA, mB])
    sc.law_p = law
    sc.law_w = scene.Law4(mode=law.mode)
    return sc, _state(sc)


@pytest.mark.parametrize("mode", [0, 1])
def test_contact_torque_closed_form(mode):
    """Normal impact at rest off A's centre: F_A = Kn*delta pushing A away
    from the face, torques (c - x) x F with the contact point c of each law:
    (cpA - cpB)/2 in the reference law (Q3, .cu:677, a lever from near the
    origin), (cpA + cpB)/2 in the physical one (.cu:815-821)."""
    from mwdiem_amd import scene
    delta, Kn = 0.05, 1e4
    A, B = _tip_into_face(delta, tip=(0.3, -0.2))
    comA, comB = np.array([1.1, 0.4, 0.9]), np.array([-0.3, 0.2, -0.6])
    sc, st = _pair_scene(A, B, comA, comB, scene.Law4(Kn=Kn, en=0.5, et=0.5, mu=0.3, mode=mode), mA=2e-3, mB=3e-3)
    coll, depth, nrm, ca, cb, F, T, est = O.contact_state(np.array([[1, 2]]), st, sc)
    assert coll[0] == 1
    cpA, cpB = A[0], np.array([A[0, 0], A[0, 1], 0.0])
    d = depth[0]  # EPA's depth, within 1e-5 of delta
    assert abs(d - delta) <= 1e-5 * delta
    Fa = np.array([0.0, 0.0, Kn * d])
    c = (cpA - cpB) / 2 if mode == 0 else (cpA + cpB) / 2
    assert np.allclose(F[1], Fa, rtol=1e-12, atol=1e-9) and np.allclose(F[2], -Fa, rtol=1e-12, atol=1e-9)
    assert np.allclose(T[1], np.cross(c - comA, Fa), rtol=1e-10, atol=1e-9)
    assert np.allclose(T[2], np.cross(c - comB, -Fa), rtol=1e-10, atol=1e-9)
    # the two laws' levers differ: the reference's contact point is near the origin
    assert np.linalg.norm(c) < 0.05 if mode == 0 else np.linalg.norm(c - cpA) < 0.05


def test_epa_face_inside_test_is_strict():
    """AddTriangle keeps a face only if the origin's projection lies strictly
    inside all three edges (.cu:300-318). With the tip straight above the
    face's centre, the projection falls on the diagonal that splits the
    Minkowski face into two triangles, both are set aside (distance 1e30),
    and EPA ends on the far face: depth 2 - delta along +z instead of delta."""
    delta = 0.05
    A, B = _tip_into_face(delta, tip=(0.0, 0.0))
    c, depth, n, cpA, cpB, st = O.gjk_epa(A, B, A.mean(0), B.mean(0))
    assert c == 1 and st == 0
    assert abs(depth - (2.0 - delta)) <= 1e-12 and np.allclose(n, [0, 0, 1], atol=1e-12)
    A2, B2 = _tip_into_face(delta, tip=(1e-3, 0.0))
    assert abs(O.gjk_epa(A2, B2, A2.mean(0), B2.mean(0))[1] - delta) <= 1e-5 * delta


def _ref_tangential_scene(vA=(0.0, 0.0, 0.0)):
    """The tip contact with hand-picked numbers for the reference law's
    tangential coefficients (.cu:685-701): mu = 1/2, en = et = exp(-pi) (so
    ln en = ln et = -pi and the two restitution factors are equal), the
    contact point c = (cpA - cpB)/2 = (0, 0, -delta/2), centres 2 above and
    3 below it and inertia sizes 8 and 18, so that
    C = 1/mu + |c - xA|^2/IA + |c - xB|^2/IB = 2 + 4/8 + 9/18 = 3,
    Kt = Kn (pi^2 + ln^2 et)/(pi^2 + ln^2 en) / C = Kn/3 and
    Gt = -2 ln(et)/C = 2 pi/3."""
    from mwdiem_amd import scene
    delta, Kn = 0.05, 1e4
    A, B = _tip_into_face(delta, tip=(0.3, -0.2))
    c_ref = np.array([0.0, 0.0, -delta / 2])
    comA, comB = c_ref + [0, 0, 2.0], c_ref - [0, 0, 3.0]
    e = math.exp(-math.pi)
    sc, st = _pair_scene(A, B, comA, comB, scene.Law4(Kn=Kn, en=e, et=e, mu=0.5, mode=0), isA=8.0, isB=18.0)
    st["v"][1] = vA
    return sc, st, Kn, delta


def test_reference_law_tangential_damping_hand_computed():
    # sliding along x at 10 mm/s: Ft = -Gt u_t = -(2 pi/3)*10, under the
    # Coulomb cap mu*Kn*delta = 250; only Fn acts on the linear momentum (Q5),
    # Ft enters the torques: tau_A = (0,0,-2) x (Ft,0,Fn) = (0, -2 Ft, 0),
    # tau_B = (0,0,3) x -(Ft,0,Fn) = (0, -3 Ft, 0)
    sc, st, Kn, delta = _ref_tangential_scene(vA=(10.0, 0.0, 0.0))
    coll, depth, *_r, F, T, est = O.contact_state(np.array([[1, 2]]), st, sc)
    assert coll[0] == 1
    d = depth[0]
    Ft = -(2 * math.pi / 3) * 10.0
    assert np.allclose(F[1], [0, 0, Kn * d], rtol=1e-12, atol=1e-9)
    assert np.allclose(T[1], [0, -2 * Ft, 0], rtol=1e-12, atol=1e-9)
    assert np.allclose(T[2], [0, -3 * Ft, 0], rtol=1e-12, atol=1e-9)


def test_reference_law_tangential_stiffness_hand_computed():
    # at rest with a stored tangential displacement xi = 0.01 along x (the
    # persistent history): Ft = -Kt xi = -(Kn/3)*0.01, under the cap; it sticks,
    # so the stored displacement comes back as xi (q_A = 1, u_t = 0)
    sc, st, Kn, delta = _ref_tangential_scene()
    xi = 0.01
    coll, depth, *_r, F, T, est, new = O.contact_state(np.array([[1, 2]]), st, sc,
                                                       dispt=np.array([[xi, 0, 0, 0]]), with_disp=True)
    Ft = -(Kn / 3) * xi
    assert np.allclose(T[1], [0, -2 * Ft, 0], rtol=1e-12, atol=1e-9)
    assert np.allclose(new[0, :3], [xi, 0, 0], rtol=1e-14, atol=0)


def test_reference_law_coulomb_cap_and_slip_displacement():
    # sliding at 1000 mm/s: Gt u_t = 2094 > mu Kn delta, so Ft = -mu |Fn| t;
    # the stored slip displacement is the reference's (-mu|Fn| t - Gt u_t)/Kt
    # (.cu:800-806 as written, the sign noted as Q19)
    sc, st, Kn, delta = _ref_tangential_scene(vA=(1000.0, 0.0, 0.0))
    coll, depth, *_r, F, T, est, new = O.contact_state(np.array([[1, 2]]), st, sc, with_disp=True)
    d = depth[0]
    Ft = -0.5 * Kn * d
    assert np.allclose(T[1], [0, -2 * Ft, 0], rtol=1e-12, atol=1e-9)
    slip = (-0.5 * Kn * d - (2 * math.pi / 3) * 1000.0) / (Kn / 3)
    assert np.allclose(new[0, :3], [slip, 0, 0], rtol=1e-12, atol=0)


# ------------------------------------------------------------ Verlet torque and gyroscopic path
def _qmul(a, b):
    w1, v1 = a[0], np.asarray(a[1:])
    w2, v2 = b[0], np.asarray(b[1:])
    return np.concatenate([[w1 * w2 - v1 @ v2], w1 * v2 + w2 * v1 + np.cross(v1, v2)])


def _axis_quat(axis, angle):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    return np.concatenate([[math.cos(angle / 2)], math.sin(angle / 2) * axis])


def test_verlet_constant_torque_about_a_principal_axis():
    """From rest, a constant world torque along the body's z axis (body turned
    by q0, so body and world frames differ): alpha = tau/Iz, and the
    staggered scheme gives w12bf = k alpha dt e_z, wbf = (k + 1/2) alpha dt e_z,
    q12 = q0 (x) rot_z(alpha dt^2 (k^2-1)/2) and
    q = q0 (x) rot_z(alpha dt^2 (k^2 + k - 3/4)/2) after k steps
    (gpu_Verlet.cu:171-271: tau_bf = q^-1 tau q, alpha = I^-1 (tau_bf - w x I w),
    the half-step turns use the previous step's world w)."""
    dt, Iz, tau = 1e-3, 3.0, 0.6
    sc = _Sc(1, dt=dt, I=(1.0, 2.0, Iz))
    st = _state(sc)
    q0 = _axis_quat([0.3, -0.5, 0.8], 1.1)
    for k in ("q", "q12", "qall"):
        st[k][1] = q0
    R0 = _rotmat(q0)
    T = np.zeros((2, 3))
    T[1] = R0 @ np.array([0.0, 0.0, tau])
    F = np.zeros((2, 3))
    alpha = tau / Iz
    for k in range(1, 41):
        O.verlet(st, F, T, sc, 1, 2)
        assert np.allclose(st["w12bf"][1], [0, 0, k * alpha * dt], rtol=1e-12, atol=1e-15)
        assert np.allclose(st["wbf"][1], [0, 0, (k + 0.5) * alpha * dt], rtol=1e-12, atol=1e-15)
        q12 = _qmul(q0, _axis_quat([0, 0, 1], alpha * dt * dt * (k * k - 1) / 2))
        q = _qmul(q0, _axis_quat([0, 0, 1], alpha * dt * dt * (k * k + k - 0.75) / 2))
        assert np.allclose(st["q12"][1], q12, rtol=0, atol=1e-14), k
        assert np.allclose(st["q"][1], q, rtol=0, atol=1e-14), k
        assert np.allclose(st["w"][1], R0 @ st["wbf"][1], rtol=1e-12, atol=1e-15)


def test_verlet_torque_free_asymmetric_body_conserves_momentum_and_energy():
    """A torque-free body with three different moments (full tensor, not
    diagonal in the body frame) spinning off its principal axes: the world
    angular momentum R(q) I w_bf and the energy w_bf.I w_bf / 2 stay constant
    to the scheme's O(dt^2); without the gyroscopic term w x I w, or with its
    sign or I and I^-1 swapped, the momentum turns or the energy drifts by
    O(1)."""
    rng = np.random.default_rng(2)
    Rb = _rotmat(rng.normal(size=4))
    I = Rb @ np.diag([1.0, 2.5, 4.0]) @ Rb.T
    dt = 2e-4
    sc = _Sc(1, dt=dt)
    sc.I[1] = I
    sc.Iinv[1] = np.linalg.inv(I)
    st = _state(sc)
    q0 = _axis_quat([1.0, 2.0, -0.5], 0.7)
    wbf = np.array([3.0, -1.5, 2.0])
    for k in ("q", "q12", "qall"):
        st[k][1] = q0
    st["wbf"][1] = st["w12bf"][1] = wbf
    st["w"][1] = _rotmat(q0) @ wbf
    F = np.zeros((2, 3))
    L0 = _rotmat(q0) @ I @ wbf
    E0 = 0.5 * wbf @ I @ wbf
    for _ in range(10000):  # 2 s, about 8 turns
        O.verlet(st, F, F, sc, 1, 2)
    q, w = st["q"][1], st["wbf"][1]
    L = _rotmat(q) @ I @ w
    E = 0.5 * w @ I @ w
    assert np.linalg.norm(L - L0) / np.linalg.norm(L0) < 1e-4
    assert abs(E / E0 - 1) < 1e-4
    # the body did turn, and not about a fixed body axis
    assert np.linalg.norm(w - wbf) > 0.5


def test_oracle_results_do_not_depend_on_its_thread_count():
    """The oracle evaluates pairs and particles in parallel (OpenMP) and adds
    contact sums in pair order: one thread and all threads give the same
    state bit for bit after steps with contacts."""
    import hashlib
    import os
    import subprocess
    import sys
    code = (
        "import hashlib, numpy as np\n"
        "from mwdiem_amd import scene\n"
        "from oracle import oracle as O\n"
        "sc = scene.compact_dense(scene.make_scene(40, (8.0, 8.0, 16.0), 7, law=scene.Law4(mode=1)), gap=-0.55)\n"
        "st = O.initial_state(sc)\n"
        "n = 0\n"
        "for _ in range(5):\n"
        "    n += int((O.step(st, sc)[1] == 1).sum())\n"
        "h = hashlib.sha256(b''.join(st[k].tobytes() for k in O.STATE_KEYS)).hexdigest()\n"
        "print(n, h)\n")
    root = Path(__file__).resolve().parents[1]
    outs = []
    for threads in ("1", "4"):
        env = dict(os.environ, OMP_NUM_THREADS=threads, PYTHONPATH=str(root))
        r = subprocess.run([sys.executable, "-c", code], env=env, cwd=root, capture_output=True, text=True,
                           check=True)
        outs.append(r.stdout.split())
    assert int(outs[0][0]) > 20 and outs[0] == outs[1]

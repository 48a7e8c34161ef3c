"""TEST INFRASTRUCTURE ONLY -- Python face of the oracle.

Loads oracle/libmwdiem_oracle.so (a plain g++ build of mwdiem_oracle.cpp, the
FP64 CPU restatement of the reference's per-timestep path) and adds a
pure-Python restatement of the reference's Linked_cell_hash that uses real
CPython dicts and sets. Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may use this module.

Parity status of each function (what pins it; see tests/test_oracle_*.py):
  cell_keys ........... closed form (truncation, linear key) + numpy
  wall_pairs .......... brute force numpy slab test
  linked_cell_hash .... pure-Python dict/set restatement (this file) and an
                        O(N^2) brute-force multiset of (cell offset) pairs
  set_order ........... CPython's own set iteration order
  contact (GJK/EPA) ... closed forms (overlapping cubes: depth, normal,
                        contact slab), separated pairs, A/B swap symmetry,
                        Minkowski-difference hull distance (scipy Qhull)
  force law ........... hand-derived spring-dashpot cases (normal and
                        Coulomb-capped tangential), Newton's third law
  verlet .............. free fall closed form, torque-free spin closed form,
                        |q| = 1, out-of-box reset, quaternion KATs
                        (test.py:16-31 -> (-60,12,30,24))
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

# MWD_ORACLE_LIB: another build of the oracle (for checking that the pins
# catch planted mistakes in a modified copy of the source)
LIB = Path(os.environ.get("MWD_ORACLE_LIB") or Path(__file__).resolve().parent / "libmwdiem_oracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise OSError(f"{LIB} not built (make -C oracle)")
        _lib = C.CDLL(str(LIB))
        _lib.orc_wall_pairs.restype = C.c_int64
        _lib.orc_linked_cell_hash.restype = C.c_int64
        _lib.orc_count_flops.restype = C.c_uint64
        _lib.orc_contact_state.restype = C.c_int64
        _lib.orc_epa_max_faces.restype = C.c_int
        _lib.orc_gjk_epa.restype = C.c_int
        _lib.orc_pair_flops.restype = C.c_uint64
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def set_order(ids) -> list:
    ids = _i64(ids)
    out = np.zeros_like(ids)
    lib().orc_set_order(_p(ids), C.c_int64(ids.size), _p(out))
    return out.tolist()


def cell_keys(com, cell, ncx, ncy) -> np.ndarray:
    com = _f64(com)
    keys = np.zeros(com.shape[0], dtype=np.int64)
    lib().orc_cell_keys(_p(com), C.c_int64(com.shape[0]), C.c_double(cell[0]), C.c_double(cell[1]),
                        C.c_double(cell[2]), C.c_int64(ncx), C.c_int64(ncy), _p(keys))
    return keys


def wall_pairs(com, W, n, wall_n, wall_c, thr) -> np.ndarray:
    com, wall_n, wall_c = _f64(com), _f64(wall_n), _f64(wall_c)
    cnt = lib().orc_wall_pairs(_p(com), C.c_int64(W), C.c_int64(n), _p(wall_n), _p(wall_c), C.c_double(thr), None,
                               C.c_int64(0))
    out = np.zeros((cnt, 2), dtype=np.int32)
    lib().orc_wall_pairs(_p(com), C.c_int64(W), C.c_int64(n), _p(wall_n), _p(wall_c), C.c_double(thr), _p(out),
                         C.c_int64(cnt))
    return out


def linked_cell_hash(keys, pmin, pmax, ncx, ncy, stencil: int = 0) -> np.ndarray:
    """stencil 0: the reference's 7 offsets; 1: the full 13-cell half shell."""
    keys = _i64(keys)
    cnt = lib().orc_linked_cell_hash(_p(keys), C.c_int64(pmin), C.c_int64(pmax), C.c_int64(ncx), C.c_int64(ncy),
                                     None, C.c_int64(0), C.c_int(stencil))
    out = np.zeros((cnt, 2), dtype=np.int32)
    lib().orc_linked_cell_hash(_p(keys), C.c_int64(pmin), C.c_int64(pmax), C.c_int64(ncx), C.c_int64(ncy), _p(out),
                               C.c_int64(cnt), C.c_int(stencil))
    return out


def linked_cell_hash_py(keys, pmin, pmax, ncx, ncy) -> list:
    """Linked_cell_hash (LIBRARY_Coarse_contact_detection.py:84-155) restated
    with real CPython dict/set objects, so dict insertion order and set
    iteration order are the interpreter's own."""
    hashdict = {}
    for i in range(pmin, pmax):
        hashdict[i] = int(keys[i])
    rev = {}
    for k, v in hashdict.items():
        rev.setdefault(v, set()).add(k)
    out = []
    offs = [1, ncx, ncx * ncy, ncx * ncy + 1, ncx * ncy + 1 + ncx, ncx * ncy + ncx, ncx + 1]
    for key in rev:
        L = list(rev[key])
        for j in range(len(L)):
            for k in range(j + 1, len(L)):
                out.append([L[j], L[k]])
        for o in offs:
            if key + o in rev:
                M = list(rev[key + o])
                for a in L:
                    for b in M:
                        out.append([a, b])
    return out


def law(Kn, en, et, mu, dt) -> np.ndarray:
    out = np.zeros(13)
    lib().orc_law(C.c_double(Kn), C.c_double(en), C.c_double(et), C.c_double(mu), C.c_double(dt), _p(out))
    return out


def contact_pairs(comA, comB, vertsA, vertsB, vA, vB, wA, wB, nA, nB, idA, idB, dispt, qA, qB, isA, isB,
                  law_p, law_w, W, dt, force=None, torque=None):
    """Per-pair fine contact in the reference's per-pair layout. Returns a dict
    of per-pair outputs; force/torque (if given) are updated in place."""
    P = comA.shape[0]
    vmax = vertsA.shape[1]
    a = [_f64(comA), _f64(comB), _f64(vertsA), _f64(vertsB), _f64(vA), _f64(vB), _f64(wA), _f64(wB)]
    ia = [_i32(nA), _i32(nB), _i32(idA), _i32(idB)]
    b = [_f64(dispt), _f64(qA), _f64(qB), _f64(isA), _f64(isB)]
    lp = _f64(list(law_p) + [0.0] * (5 - len(law_p)))
    lw = _f64(list(law_w) + [0.0] * (5 - len(law_w)))
    out = dict(collision=np.zeros(P, np.int32), epa_status=np.zeros(P, np.int32), depth=np.zeros(P),
               normal=np.zeros((P, 3)), cpA=np.zeros((P, 3)), cpB=np.zeros((P, 3)), Fn=np.zeros((P, 3)),
               tauA=np.zeros((P, 3)), tauB=np.zeros((P, 3)), disp=np.zeros((P, 4)), faces=np.zeros(P, np.int32))
    lib().orc_contact_pairs(C.c_int64(P), C.c_int(vmax), *[_p(x) for x in a], *[_p(x) for x in ia],
                            *[_p(x) for x in b], _p(lp), _p(lw), C.c_int64(W), C.c_double(dt),
                            _p(out["collision"]), _p(out["epa_status"]), _p(out["depth"]), _p(out["normal"]),
                            _p(out["cpA"]), _p(out["cpB"]), _p(out["Fn"]), _p(out["tauA"]), _p(out["tauB"]),
                            _p(out["disp"]), _p(force), _p(torque), _p(out["faces"]))
    return out


def _law5(L):
    return np.array([L.Kn, L.en, L.et, L.mu, float(getattr(L, "mode", 0))])


def _laws(sc):
    return _law5(sc.law_p), _law5(sc.law_w)


def contact_state(pairs, st, sc, dispt=None, with_disp=False):
    """Fine contact over a pair list indexed into state arrays; returns
    (collision, depth, normal, cpA, cpB, force_sum, torque_sum, epa_status),
    and the pairs' new tangential displacements [P,4] with with_disp. dispt
    [P,4]: the pairs' tangential history (None: zero, the reference's reset)."""
    pairs = _i32(pairs).reshape(-1, 2)
    P = pairs.shape[0]
    pa, pb = np.ascontiguousarray(pairs[:, 0]), np.ascontiguousarray(pairs[:, 1])
    N = st["com"].shape[0]
    coll = np.zeros(P, np.int32)
    est = np.zeros(P, np.int32)
    depth = np.zeros(P)
    nrm = np.zeros((P, 3))
    ca = np.zeros((P, 3))
    cb = np.zeros((P, 3))
    F = np.zeros((N, 3))
    T = np.zeros((N, 3))
    dispt = None if dispt is None else _f64(dispt)
    disp = np.zeros((P, 4)) if with_disp else None
    lp, lw = _laws(sc)
    lib().orc_contact_state(C.c_int64(P), _p(pa), _p(pb), C.c_int(sc.vmax), _p(_f64(st["com"])),
                            _p(_f64(st["verts"])), _p(_i32(sc.nverts)), _p(_f64(st["v"])), _p(_f64(st["w"])),
                            _p(_f64(st["q"])), _p(_f64(sc.isize)), _p(_f64(sc.mass)), _p(lp), _p(lw),
                            C.c_int64(sc.n_walls), C.c_double(sc.dt), _p(coll), _p(depth), _p(nrm), _p(ca), _p(cb), _p(F), _p(T), _p(est),
                            _p(dispt), _p(disp))
    if disp is not None:
        return coll, depth, nrm, ca, cb, F, T, est, disp
    return coll, depth, nrm, ca, cb, F, T, est


def contact_faces(pairs, st, sc):
    """Collision codes of a pair list and, for colliding pairs, EPA's final
    face: the support indices of its three Minkowski vertices on A
    (face[:, :3]) and on B (face[:, 3:]), in the face's stored order (-1 for
    the others); each pair takes the GJK/EPA variant of its law's mode."""
    pairs = _i32(pairs).reshape(-1, 2)
    P = pairs.shape[0]
    pa, pb = np.ascontiguousarray(pairs[:, 0]), np.ascontiguousarray(pairs[:, 1])
    coll = np.zeros(P, np.int32)
    face = np.zeros((P, 6), np.int32)
    lib().orc_contact_faces(C.c_int64(P), _p(pa), _p(pb), C.c_int(sc.vmax), _p(_f64(st["com"])),
                            _p(_f64(st["verts"])), _p(_i32(sc.nverts)), _p(coll), _p(face),
                            C.c_int(int(getattr(getattr(sc, "law_p", None), "mode", 0))),
                            C.c_int(int(getattr(getattr(sc, "law_w", None), "mode", 0))),
                            C.c_int64(int(getattr(sc, "n_walls", 0))))
    return coll, face


def gjk_trace(pairs, com, verts, nverts, robust: bool = False):
    """Path of GJK/EPA for each pair of a list: (collision, tetrahedron
    iterations, CorrectFlat fired (0/1), EPA status, EPA expansions).
    robust: the physical law's variant (flat tetrahedra re-supported along
    the face normal, EPA faces turned away from an interior point)."""
    pairs = _i32(pairs).reshape(-1, 2)
    P = pairs.shape[0]
    pa, pb = np.ascontiguousarray(pairs[:, 0]), np.ascontiguousarray(pairs[:, 1])
    verts = _f64(verts)
    out = [np.zeros(P, np.int32) for _ in range(5)]
    lib().orc_gjk_trace(C.c_int64(P), _p(pa), _p(pb), C.c_int(verts.shape[1]), _p(_f64(com)), _p(verts),
                        _p(_i32(nverts)), *[_p(x) for x in out], C.c_int(1 if robust else 0))
    return tuple(out)


def count_flops(pairs, st, sc) -> int:
    pairs = _i32(pairs).reshape(-1, 2)
    pa, pb = np.ascontiguousarray(pairs[:, 0]), np.ascontiguousarray(pairs[:, 1])
    lp, lw = _laws(sc)
    return int(lib().orc_count_flops(C.c_int64(pairs.shape[0]), _p(pa), _p(pb), C.c_int(sc.vmax),
                                     _p(_f64(st["com"])), _p(_f64(st["verts"])), _p(_i32(sc.nverts)),
                                     _p(_f64(st["v"])), _p(_f64(st["w"])), _p(_f64(st["q"])), _p(_f64(sc.isize)),
                                     _p(_f64(sc.mass)), _p(lp), _p(lw), C.c_int64(sc.n_walls), C.c_double(sc.dt)))


STATE_KEYS = ("com", "v", "v12", "w", "wbf", "w12bf", "q", "q12", "qall", "dxall", "verts")


def initial_state(sc) -> dict:
    """The driver's initial arrays (MWDiem.py:603-647): at rest, identity quaternions."""
    N = sc.n_rows
    st = {k: np.zeros((N, 3)) for k in ("v", "v12", "w", "wbf", "w12bf", "dxall")}
    for k in ("q", "q12", "qall"):
        st[k] = np.zeros((N, 4))
        st[k][:, 0] = 1.0
    st["com"] = sc.com.copy()
    st["verts"] = sc.verts.copy()
    if getattr(sc, "wall_v", None) is not None:  # prescribed wall velocities
        st["v"][:sc.n_walls] = sc.wall_v[:sc.n_walls]
        st["wall_c"] = sc.wall_c.copy()
    return st


def verlet(st, F, T, sc, pmin, pmax, nrot=None, rot_mode=None):
    """Verlet (gpu_Verlet.cu:85-404) in place on the state dict. rot_mode 0
    rotates vertices by q (reference), 1 by the increment q q_old^-1; by
    default it follows the scene's particle law mode."""
    if rot_mode is None:
        rot_mode = int(getattr(sc.law_p, "mode", 0))
    arr = {k: st[k] for k in STATE_KEYS}
    for k, a in arr.items():
        assert a.dtype == np.float64 and a.flags.c_contiguous, k
    lib().orc_verlet(_p(arr["v"]), _p(arr["v12"]), _p(arr["w12bf"]), _p(arr["q12"]), _p(arr["q"]), _p(arr["qall"]),
                     _p(arr["wbf"]), _p(arr["w"]), _p(_f64(F)), _p(_f64(T)), _p(arr["verts"]), _p(arr["com"]),
                     _p(_f64(sc.mass)), _p(_f64(sc.Iinv)), _p(_f64(sc.I)), _p(arr["dxall"]), C.c_int(sc.vmax),
                     _p(None if nrot is None else _i32(nrot)), C.c_double(sc.dt), _p(_f64(sc.box)),
                     C.c_int64(pmin), C.c_int64(pmax), C.c_int(rot_mode))


def candidates(st, sc, n_active: int | None = None) -> np.ndarray:
    """Wall slab pairs then Linked_cell_hash pairs (MWDiem.py:807-817) over
    the active rows [n_walls, n_active) (num_particles_temp; all rows by
    default)."""
    W = sc.n_walls
    N = sc.n_rows if n_active is None else n_active
    wp = wall_pairs(st["com"], W, N, sc.wall_n, st.get("wall_c", sc.wall_c), sc.wall_thr)
    keys = cell_keys(st["com"], sc.cell, sc.ncx, sc.ncy)
    hp = linked_cell_hash(keys, W, N, sc.ncx, sc.ncy, stencil=int(getattr(sc.law_p, "mode", 0)))
    return np.concatenate([wp, hp]).astype(np.int32).reshape(-1, 2)


def pair_keys(pairs) -> np.ndarray:
    """The driver's tangential-history key min + max^2 (MWDiem.py:872-880)."""
    a = np.minimum(pairs[:, 0], pairs[:, 1]).astype(np.int64)
    b = np.maximum(pairs[:, 0], pairs[:, 1]).astype(np.int64)
    return a + b * b


def step(st, sc, history: dict | None = None, n_active: int | None = None):
    """One iteration of the reference time loop (MWDiem.py:805-1001) in place;
    returns (pairs, collision). With `history` (a dict key -> [4]), the
    tangential history persists: each pair reads its entry and the dict is
    replaced by this step's contacts' new displacements (the history the
    driver intends; as it runs, it drops every row, Q1). n_active: rows
    [n_walls, n_active) take part (num_particles_temp, MWDiem.py:685-714,
    808-817, 1001); all rows by default."""
    N = sc.n_rows if n_active is None else n_active
    pairs = candidates(st, sc, N)
    if history is None:
        coll, _, _, _, _, F, T, _ = contact_state(pairs, st, sc)
    else:
        keys = pair_keys(pairs)
        dispt = np.zeros((pairs.shape[0], 4))
        for i, k in enumerate(keys.tolist()):
            if k in history:
                dispt[i] = history[k]
        coll, _, _, _, _, F, T, _, disp = contact_state(pairs, st, sc, dispt=dispt, with_disp=True)
        history.clear()
        for i in np.flatnonzero(coll == 1).tolist():
            history.setdefault(int(keys[i]), disp[i].copy())
    F = F + sc.fext
    T = T + sc.text
    verlet(st, F, T, sc, sc.n_walls, N, nrot=sc.nverts)
    if getattr(sc, "wall_v", None) is not None:
        move_walls(st, sc)
    return pairs, coll


def move_walls(st, sc):
    """Walls with a prescribed velocity translate by v dt at the end of the
    step: centre, vertices, slab centre (a kinematic boundary; the
    reference's walls are static)."""
    W = sc.n_walls
    d = st["v"][:W] * sc.dt
    st["com"][:W] += d
    if "wall_c" in st:
        st["wall_c"][:W] += d
    for r in range(W):
        st["verts"][r, :int(sc.nverts[r])] += d[r]


def quat_mul(a, b) -> np.ndarray:
    out = np.zeros(4)
    lib().orc_quat_mul(_p(_f64(a)), _p(_f64(b)), _p(out))
    return out


def quat_tripl(q1, v, q2) -> np.ndarray:
    out = np.zeros(3)
    lib().orc_quat_tripl(_p(_f64(q1)), _p(_f64(v)), _p(_f64(q2)), _p(out))
    return out


def gjk_epa(VA, VB, comA, comB, robust: bool = False):
    """(collision, depth, normal, cpA, cpB, epa_status) of two vertex sets;
    robust: the physical law's GJK/EPA variant."""
    VA, VB = _f64(VA), _f64(VB)
    out = np.zeros(10)
    st = C.c_int(0)
    c = lib().orc_gjk_epa(_p(VA), C.c_int(VA.shape[0]), _p(VB), C.c_int(VB.shape[0]), _p(_f64(comA)),
                          _p(_f64(comB)), _p(out), C.byref(st), C.c_int(1 if robust else 0))
    return c, out[0], out[1:4], out[4:7], out[7:10], st.value


def pair_flops(VA, VB, comA, comB) -> int:
    VA, VB = _f64(VA), _f64(VB)
    return int(lib().orc_pair_flops(_p(VA), C.c_int(VA.shape[0]), _p(VB), C.c_int(VB.shape[0]), _p(_f64(comA)),
                                    _p(_f64(comB))))

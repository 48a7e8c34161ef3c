"""TEST INFRASTRUCTURE ONLY -- the reference's CPU step structure in pure
Python, for the CPU baseline.

A Python restatement of one iteration of the reference time loop
(MWDiem.py:805-1001) with the reference's own host structure: cell keys with
numpy (MWDiem.py:813-816); the wall slab test as a Python double loop over
walls and particles with a numpy dot per test (:807-811); Linked_cell_hash
with CPython dicts and sets (LIBRARY_Coarse_contact_detection.py:84-155,
oracle.linked_cell_hash_py); the per-pair gather loop into per-pair arrays
(MWDiem.py:825-903); the fine contact of every candidate pair as scalar
Python -- GJK, EPA and the contact law of gpu_fine_contact_detection.cu:41-871
executed one pair after another on one core, in FP64; the force assembly with
numpy (:977-981); and the Verlet step one particle after another
(gpu_Verlet.cu:85-404). It computes what oracle/mwdiem_oracle.cpp computes
(tests/test_ref_structure.py checks them equal on small scenes); it exists
to time the reference's structure on the host, and only bench.py's
cpu_baseline leg and tests/ use it.

Usage (timing): python -m oracle.ref_structure --config c1|c2 --steps N
"""
from __future__ import annotations

import math

import numpy as np

from . import oracle as O

EPA_MAX_FACES = 128
EPA_PLACEHOLDER = 1e20


# ------------------------------------------------------------ vectors (tuples)
def add(a, b):
    return (a[0] + b[0], a[1] + b[1], a[2] + b[2])


def sub(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def mul(a, s):
    return (a[0] * s, a[1] * s, a[2] * s)


def div(a, s):
    return (a[0] / s, a[1] / s, a[2] / s)


def neg(a):
    return (-a[0], -a[1], -a[2])


def dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def norm2(a):
    return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]


def is_zero(a):
    return a[0] == 0.0 and a[1] == 0.0 and a[2] == 0.0


def qmul(a, b):
    """Quat_mult (gpu_Verlet.cu:43-54), (w, x, y, z)."""
    v1, v2 = a[1:], b[1:]
    s = add(add(mul(v2, a[0]), mul(v1, b[0])), cross(v1, v2))
    return (a[0] * b[0] - dot(v1, v2),) + s


def qinv(q):
    return (q[0], -q[1], -q[2], -q[3])


def qtripl(q1, v, q2):
    """Quat_tripl (gpu_Verlet.cu:57-74): vector part of q1 (0,v) q2."""
    return qmul(q1, qmul((0.0,) + tuple(v), q2))[1:]


# ------------------------------------------------------------ GJK (.cu:55-254)
def support(A, B, d):
    """Support_funtion (.cu:55-101): first strict maximum on each body."""
    best, i1 = -1e30, 0
    for i, a in enumerate(A):
        t = dot(d, a)
        if t > best:
            best, i1 = t, i
    nd = neg(d)
    best2, i2 = -1e30, 0
    for i, b in enumerate(B):
        t = dot(nd, b)
        if t > best2:
            best2, i2 = t, i
    return sub(A[i1], B[i2]), i1, i2


def check_segment(P1, P2):
    """Check_linesegment (.cu:103-124): (coll, d)."""
    V21 = sub(P2, P1)
    if dot(V21, P2) < 0.0:
        return 0, (0.0, 0.0, 0.0)
    if is_zero(cross(V21, P2)):
        return 1, (0.0, 0.0, 0.0)
    return -1, cross(cross(V21, neg(P2)), V21)


def check_triangle(P1, P2, P3):
    """Check_triangle (.cu:127-186)."""
    c2, d = check_segment(P2, P3)
    if c2 == 1:
        return 1, d
    c3, d = check_segment(P1, P3)
    if c3 == 1:
        return 1, d
    if c2 + c3 > -1:
        return 0, d
    V32, V13 = sub(P3, P2), sub(P1, P3)
    n = cross(V32, V13)
    n2, n3 = cross(V32, n), cross(V13, n)
    if dot(n2, P2) < 0.0:
        return -1, n2
    if dot(n3, P3) < 0.0:
        return -1, n3
    if dot(P3, n) == 0.0:
        return 1, d
    return -1, (neg(n) if dot(n, P3) > 0.0 else n)


def check_tetra(P1, P2, P3, P4):
    """Check_Tetrahedron (.cu:188-254): (coll, d, todel)."""
    c1, d = check_triangle(P1, P2, P4)
    if c1 == 1:
        return 1, d, 0
    c2, d = check_triangle(P2, P3, P4)
    if c2 == 1:
        return 1, d, 0
    c3, d = check_triangle(P3, P1, P4)
    if c3 == 1:
        return 1, d, 0
    if c1 + c2 + c3 > -2:
        return 0, d, 0
    V13, V21, V32, V41 = sub(P1, P3), sub(P2, P1), sub(P3, P2), sub(P4, P1)
    n2 = cross(V21, V41)
    n3 = cross(V32, sub(P4, P2))
    n4 = cross(V13, sub(P4, P3))
    if dot(n2, P4) < 0.0:
        return -1, n2, 3
    if dot(n3, P4) < 0.0:
        return -1, n3, 1
    if dot(n4, P4) < 0.0:
        return -1, n4, 2
    return 1, d, 0


def correct_flat(S1, S2, S3, S4):
    """CorrectFlat (.cu:256-268)."""
    det1 = S1[0] * (S2[1] * S3[2] - S3[1] * S2[2]) - S2[0] * (S1[1] * S3[2] - S3[1] * S1[2]) \
        + S3[0] * (S1[1] * S2[2] - S2[1] * S1[2])
    if -0.000001 < det1 < 0.000001:
        det2 = S1[0] * (S2[1] * S4[2] - S4[1] * S2[2]) - S2[0] * (S1[1] * S4[2] - S4[1] * S1[2]) \
            + S4[0] * (S1[1] * S2[2] - S2[1] * S1[2])
        if -0.000001 < det2 < 0.000001:
            return True
    return False


# ------------------------------------------------------------ EPA (.cu:270-603)
def add_triangle(P, e0, e1, e2, fid):
    """AddTriangle (.cu:270-325) into face fid of the polytope dict P."""
    V = P["V"]
    nn = cross(sub(V[e1], V[e0]), sub(V[e2], V[e1]))
    nn = div(nn, math.sqrt(norm2(nn)))
    dd = dot(nn, V[e0])
    if dd < 0.0:
        e1, e2 = e2, e1
        dd = dd * -1.0
        nn = mul(nn, -1.0)
    tri = (e0, e1, e2)
    a, b, c = V[e0], V[e1], V[e2]
    inside = dot(cross(sub(b, a), nn), neg(a)) < 0.0 and dot(cross(sub(c, b), nn), neg(b)) < 0.0 and \
        dot(cross(sub(a, c), nn), neg(c)) < 0.0
    valid = 1
    if not inside:
        valid, dd = -2, 1e30
    face = (tri, nn, dd, valid)
    if fid == len(P["F"]):
        P["F"].append(face)
    else:
        P["F"][fid] = face


def epa(S, ia, ib, A, B):
    """EPA (.cu:328-603): (depth, normal, cpA, cpB, status)."""
    P = {"V": list(S), "opA": list(ia), "opB": list(ib), "F": []}
    xmin_old = 1e29
    proj = 2.0 * xmin_old + 1.0
    counter, st, indmin, xmin_new = 0, 0, 0, 0.0
    for fid, (e0, e1, e2) in enumerate(((0, 1, 2), (1, 2, 3), (2, 0, 3), (0, 1, 3))):
        add_triangle(P, e0, e1, e2, fid)
    F = P["F"]
    while xmin_old < proj:
        indmin, xm = 0, F[0][2]
        for k in range(len(F)):
            if F[k][2] < xm:
                xm, indmin = F[k][2], k
        xmin_new = F[indmin][2]
        xmin_old = xmin_new
        nrm = F[indmin][1]
        Suj, a, b = support(A, B, nrm)
        proj = dot(nrm, Suj)
        if xmin_old > proj or (xmin_old * 1.00001 > proj and xmin_old * 0.99999 < proj):
            break
        if counter > 49:
            st |= 1
            break
        counter += 1
        P["V"].append(Suj)
        P["opA"].append(a)
        P["opB"].append(b)
        nv = len(P["V"])
        V = P["V"]
        # faces that see the new point (valid or outside-placeholder), .cu:461-467
        vis = []
        for i in range(len(F)):
            tri, n, dd, valid = F[i]
            if valid in (1, -2) and dot(sub(V[nv - 1], V[tri[0]]), n) >= 0.0:
                F[i] = (tri, n, dd, -1)
                vis.append(i)
        # horizon: edges of visible faces no other visible face shares (.cu:470-540)
        count = {}
        for i in vis:
            t = F[i][0]
            for e in range(3):
                u, w = t[e], t[(e + 1) % 3]
                key = (u, w) if u < w else (w, u)
                count[key] = count.get(key, 0) + 1
        overflow = False
        for i in vis:
            t = F[i][0]
            for e in range(3):
                u, w = t[e], t[(e + 1) % 3]
                if count[(u, w) if u < w else (w, u)] >= 2:
                    continue
                if len(F) == EPA_MAX_FACES:
                    overflow = True
                    break
                add_triangle(P, u, w, nv - 1, len(F))
            if overflow:
                break
            tri, n, _, _ = F[i]
            F[i] = (tri, n, 3e20, 0)
        if overflow:
            st |= 2
            break
        counter += 1
    depth = xmin_new
    (a, b, c), nrm, _, _ = F[indmin]
    V = P["V"]
    Proj = mul(nrm, depth)
    Aa, Bb, Cc = V[a], V[b], V[c]
    v0, v1, v2 = sub(Bb, Aa), sub(Cc, Aa), sub(Proj, Aa)
    if dot(cross(v0, v1), nrm) < 0.0:
        v0, v1 = sub(Cc, Aa), sub(Bb, Aa)
    v00, v01, v11, v20, v21 = dot(v0, v0), dot(v0, v1), dot(v1, v1), dot(v2, v0), dot(v2, v1)
    den = v00 * v11 - v01 * v01
    vv = (v11 * v20 - v01 * v21) / den
    ww = (v00 * v21 - v01 * v20) / den
    uu = 1.0 - ww - vv
    opA, opB = P["opA"], P["opB"]
    cpA = add(add(mul(A[opA[a]], uu), mul(A[opA[b]], vv)), mul(A[opA[c]], ww))
    cpB = add(add(mul(B[opB[a]], uu), mul(B[opB[b]], vv)), mul(B[opB[c]], ww))
    return depth, nrm, cpA, cpB, st


# ------------------------------------------------------------ Force (.cu:647-871)
def law_coeffs(L, dt):
    pi2 = 3.141592653589793 * 3.141592653589793
    ln_en, ln_et = math.log(L.en), math.log(L.et)
    A_en = pi2 + ln_en * ln_en
    tc = math.sqrt((L.mu * A_en) / L.Kn)
    return {"Kn": L.Kn, "mu": L.mu, "ln_et": ln_et, "Gn": ((-2.0 * L.mu) / tc) * ln_en,
            "Kt_num": ((L.Kn * (pi2 + ln_et * ln_et)) / A_en) * 1.0, "inv_mu": 1.0 / L.mu, "dt": dt,
            "cn": (-2.0 * ln_en) / math.sqrt(A_en), "ct": (-2.0 * ln_et) / math.sqrt(pi2 + ln_et * ln_et),
            "ratio": (pi2 + ln_et * ln_et) / A_en, "mode": int(L.mode)}


def force(pr, depth, n, cpA, cpB, L):
    """Force: (F on A's linear momentum, tauA, tauB), the history reset (Q1)."""
    phys = L["mode"] == 1
    if phys and not depth < EPA_PLACEHOLDER:
        z = (0.0, 0.0, 0.0)
        return z, z, z
    cp = div(add(cpA, cpB), 2.0) if phys else div(sub(cpA, cpB), 2.0)
    dA, dB = sub(cp, pr["comA"]), sub(cp, pr["comB"])
    dAn, dBn = norm2(dA), norm2(dB)
    if not phys:
        C = (L["inv_mu"] + dAn / pr["isA"]) + dBn / pr["isB"]
        Gt = (-2.0 / C) * L["ln_et"]
        Gn = L["Gn"]
    else:
        imA, imB = pr["imA"], pr["imB"]
        imn = imA + imB
        mn = 1.0 / imn
        imt = (imn + dAn * (0.0 if imA == 0.0 else 1.0 / pr["isA"])) + dBn * (0.0 if imB == 0.0 else 1.0 / pr["isB"])
        mt = 1.0 / imt
        Gn = L["cn"] * math.sqrt(mn * L["Kn"])
        Kt = (L["Kn"] * L["ratio"]) * (mt / mn)
        Gt = L["ct"] * math.sqrt(mt * Kt)
    if dot(sub(pr["comB"], pr["comA"]), n) < 0.0:
        n = mul(n, -1.0)
    n = div(n, math.sqrt(norm2(n)))
    disp_n = mul(n, depth)
    u = sub(sub(add(pr["vA"], cross(pr["wA"], sub(cpA, pr["comA"]))), pr["vB"]), cross(pr["wB"], sub(cpB, pr["comB"])))
    un = mul(n, dot(u, n))
    ut = sub(u, un)
    utn = math.sqrt(norm2(ut))
    t = div(ut, utn) if utn != 0.0 else (0.0, 0.0, 0.0)
    Fn = sub(mul(mul(disp_n, -1.0), L["Kn"]), mul(un, Gn))
    Fnn = math.sqrt(norm2(Fn))
    Ft = sub(mul((0.0, 0.0, 0.0), -1.0), mul(ut, Gt))  # the spring term: history zero (Q1)
    if math.sqrt(norm2(Ft)) > L["mu"] * Fnn:
        Ft = mul(mul(mul(t, Fnn), -1.0), L["mu"])
    Ff = add(Ft, Fn)
    return (Ff if phys else Fn), cross(sub(cp, pr["comA"]), Ff), cross(sub(cp, pr["comB"]), mul(Ff, -1.0))


def contact_pair(A, B, pr, L):
    """contact_detection (.cu:874-1322) for one pair: (collision, F, tauA, tauB)."""
    z = (0.0, 0.0, 0.0)
    d1 = sub(pr["comB"], pr["comA"])
    S0, a0, b0 = support(A, B, d1)
    if dot(d1, S0) < 0.0:
        return 0, z, z, z
    d2 = neg(d1)
    S1, a1, b1 = support(A, B, d2)
    if dot(d2, S1) < 0.0:
        return 0, z, z, z
    coll, d3 = check_segment(S0, S1)
    if coll == 0:
        return 0, z, z, z
    S2, a2, b2 = support(A, B, d3)
    if dot(d3, S2) < 0.0:
        return 0, z, z, z
    coll, d4 = check_triangle(S0, S1, S2)
    if coll == 0:
        return 0, z, z, z
    for _ in range(50):
        S3, a3, b3 = support(A, B, d4)
        if dot(d4, S3) < 0.0:
            return 0, z, z, z
        hit = is_zero(S3)
        if not hit:
            coll, d4, todel = check_tetra(S0, S1, S2, S3)
            if coll == 0:
                return 0, z, z, z
            if coll == 1:
                hit = True
            elif todel == 1:
                S0, a0, b0, S1, a1, b1, S2, a2, b2 = S1, a1, b1, S2, a2, b2, S3, a3, b3
            elif todel == 2:
                S0, a0, b0, S1, a1, b1, S2, a2, b2 = S2, a2, b2, S0, a0, b0, S3, a3, b3
            else:
                S2, a2, b2 = S3, a3, b3
        if hit:
            if correct_flat(S0, S1, S2, S3):
                d4 = cross(S0, S1)
                S3, a3, b3 = support(A, B, d4)
            depth, nrm, cpA, cpB, st = epa((S0, S1, S2, S3), (a0, a1, a2, a3), (b0, b1, b2, b3), A, B)
            F, tA, tB = force(pr, depth, nrm, cpA, cpB, L)
            return 1, F, tA, tB
    return -1, z, z, z


# ------------------------------------------------------------ Verlet (gpu_Verlet.cu:85-404)
def verlet_particle(st, sc, P, F, T, rot_mode):
    dt = sc.dt
    t3 = lambda k: tuple(st[k][P])  # noqa: E731
    v12, w12bf, wbf, w, com = t3("v12"), t3("w12bf"), t3("wbf"), t3("w"), t3("com")
    q, q12, dxall = tuple(st["q"][P]), tuple(st["q12"][P]), t3("dxall")
    m, I, Ii = sc.mass[P], sc.I[P], sc.Iinv[P]
    v32 = add(v12, mul(div(F, m), dt))
    tbf = qtripl(qinv(q), T, q)
    p1 = (I[0, 0] * wbf[0] + I[0, 1] * wbf[1] + I[0, 2] * wbf[2],
          I[1, 0] * wbf[0] + I[1, 1] * wbf[1] + I[1, 2] * wbf[2],
          I[2, 0] * wbf[0] + I[2, 1] * wbf[1] + I[2, 2] * wbf[2])
    p2 = sub(tbf, cross(wbf, p1))
    abf = (Ii[0, 0] * p2[0] + Ii[0, 1] * p2[1] + Ii[0, 2] * p2[2],
           Ii[1, 0] * p2[0] + Ii[1, 1] * p2[1] + Ii[1, 2] * p2[2],
           Ii[2, 0] * p2[0] + Ii[2, 1] * p2[1] + Ii[2, 2] * p2[2])
    w32bf = add(w12bf, mul(abf, dt))
    wn = math.sqrt(norm2(w))
    t11 = mul(div(w, wn), math.sin((wn * dt) / 2)) if wn != 0 else (0.0, 0.0, 0.0)
    q32 = qmul((math.cos((wn * dt) / 2),) + t11, q12)
    v12, q12, w12bf = v32, q32, w32bf
    dx = mul(v12, dt)
    dxall = add(dxall, dx)
    w34bf = add(w12bf, div(mul(abf, dt), 4.0))
    w34 = qtripl(q12, w34bf, qinv(q12))
    n34 = math.sqrt(norm2(w34))
    t1 = mul(div(w34, n34), math.sin((n34 * dt) / 4)) if n34 != 0 else (0.0, 0.0, 0.0)
    qold = q
    q = qmul((math.cos((n34 * dt) / 4),) + t1, q12)
    qall = qmul(q, qold)
    wbf = add(w12bf, div(mul(abf, dt), 2.0))
    w = qtripl(q, wbf, qinv(q))
    vel = add(v12, div(mul(F, dt), m))
    qr = qmul(q, qinv(qold)) if rot_mode == 1 else q
    qri = qinv(qr)
    nrot = int(sc.nverts[P])
    V = st["verts"][P]
    rows = [add(qtripl(qr, sub(tuple(V[i]), com), qri), com) for i in range(nrot)]
    com = add(com, dx)
    rows = [add(r, dx) for r in rows]
    b = sc.box
    if com[0] > b[0] or com[0] < b[1] or com[1] > b[2] or com[1] < b[3] or com[2] > b[4] or com[2] < b[5]:
        vel = v12 = w12bf = wbf = w = (0.0, 0.0, 0.0)
        du = mul(dx, -1.0)
        com = add(com, du)
        rows = [add(r, du) for r in rows]
    for k, val in (("v", vel), ("v12", v12), ("w12bf", w12bf), ("com", com), ("dxall", dxall), ("wbf", wbf),
                   ("w", w), ("q12", q12), ("q", q), ("qall", qall)):
        st[k][P] = val
    if nrot:
        V[:nrot] = rows


# ------------------------------------------------------------ one step (MWDiem.py:805-1001)
def step(st, sc, laws=None):
    """One iteration of the time loop in the reference's host structure, in
    place; returns the number of contacts."""
    W, N = sc.n_walls, sc.n_rows
    laws = laws or (law_coeffs(sc.law_p, sc.dt), law_coeffs(sc.law_w, sc.dt))
    com = st["com"]
    # wall slab test (MWDiem.py:807-811): walls outer, particles inner
    pairs = []
    wn, wc, thr = sc.wall_n, sc.wall_c, sc.wall_thr
    for k in range(W):
        for j in range(W, N):
            if abs(np.dot(wn[k], com[j] - wc[k])) < thr:
                pairs.append([j, k])
    # cell keys (MWDiem.py:813-816) and Linked_cell_hash with dicts and sets
    keys = (com[:, 0] / sc.cell[0]).astype(int) + (com[:, 1] / sc.cell[1]).astype(int) * sc.ncx + \
        (com[:, 2] / sc.cell[2]).astype(int) * sc.ncx * sc.ncy
    if int(sc.law_p.mode) == 0:
        pairs += O.linked_cell_hash_py(keys, W, N, sc.ncx, sc.ncy)
    else:
        pairs += O.linked_cell_hash(keys, W, N, sc.ncx, sc.ncy, stencil=1).tolist()
    # per-pair gather (MWDiem.py:825-903) into per-pair records
    verts, nv, v, w = st["verts"], sc.nverts, st["v"], st["w"]
    gathered = []
    for a, b in pairs:
        A = [tuple(x) for x in verts[a, :nv[a]]]
        B = [tuple(x) for x in verts[b, :nv[b]]]
        pr = {"comA": tuple(com[a]), "comB": tuple(com[b]), "vA": tuple(v[a]), "vB": tuple(v[b]),
              "wA": tuple(w[a]), "wB": tuple(w[b]), "isA": float(sc.isize[a]), "isB": float(sc.isize[b]),
              "imA": 0.0 if a < W else 1.0 / sc.mass[a], "imB": 0.0 if b < W else 1.0 / sc.mass[b]}
        gathered.append((a, b, A, B, pr))
    # fine contact of every pair, one after another, sums in pair order
    Fs = [(0.0, 0.0, 0.0)] * N
    Ts = [(0.0, 0.0, 0.0)] * N
    n_contacts = 0
    for a, b, A, B, pr in gathered:
        L = laws[1] if (a < W or b < W) else laws[0]
        c, F, tA, tB = contact_pair(A, B, pr, L)
        if c == 1:
            n_contacts += 1
            Fs[a] = add(Fs[a], F)
            Fs[b] = add(Fs[b], neg(F))
            Ts[a] = add(Ts[a], tA)
            Ts[b] = add(Ts[b], tB)
    # force assembly (MWDiem.py:977-981)
    F = np.array(Fs) + sc.fext
    T = np.array(Ts) + sc.text
    rot_mode = int(sc.law_p.mode)
    for P in range(W, N):
        verlet_particle(st, sc, P, tuple(F[P]), tuple(T[P]), rot_mode)
    return n_contacts


def main():
    import argparse
    import os
    import time

    from mwdiem_amd import scene
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", choices=("c1", "c2"), default="c1")
    ap.add_argument("--steps", type=int, default=10)
    args = ap.parse_args()
    sc = scene.config_small() if args.config == "c1" else scene.config_large()
    st = O.initial_state(sc)
    t0 = time.perf_counter()
    for i in range(args.steps):
        k = step(st, sc)
        print(f"step {i}: {k} contacts, {time.perf_counter() - t0:.1f} s", flush=True)
    dt = time.perf_counter() - t0
    print(f"{args.config}: {sc.n_particles} particles x {args.steps} steps in {dt:.2f} s on one core: "
          f"{sc.n_particles * args.steps / dt:.4g} particle-steps/s ({os.cpu_count()} logical CPUs)")


if __name__ == "__main__":
    main()

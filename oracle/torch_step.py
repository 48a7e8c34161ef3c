"""TEST INFRASTRUCTURE ONLY -- FP64 PyTorch-CPU restatement of the step on
the physical law, for reverse-mode reference gradients.

Only tests/ use this module. Each step takes its contact topology from the
oracle (oracle.py) on the detached state: the candidate pairs, which of them
collide, and EPA's final face (the support indices on A and B of its three
Minkowski vertices). Everything after that is a torch function of the state
and of theta = (Kn, en, mu), so reverse-mode autograd gives d outcome / d
theta over many steps, contact topology fixed per step:

  * face geometry (oracle add_triangle and the end of epa; reference
    gpu_fine_contact_detection.cu:270-325 and :567-601): n = normalise((Vb -
    Va) x (Vc - Vb)) over the face's stored vertex order, depth = n . Va,
    contact points from the barycentric coordinates of depth * n in the face;
  * the physical contact law (oracle force_law with mode 1: midpoint contact
    point, effective-mass damping, normal plus Coulomb-capped tangential
    force on the linear momentum; no tangential history, as the reference
    drops it every step, MWDiem.py:953-966);
  * the per-particle sums (F[a] += F, F[b] -= F, torques from the contact
    point) plus the external force and torque (MWDiem.py:977-981);
  * the Verlet step (gpu_Verlet.cu:164-403), vertices turned by the step's
    increment q q_old^-1 as the physical law does.

Law coefficients follow the oracle's make_law (mode 1): cn = -2 ln en /
sqrt(pi^2 + ln^2 en), ct likewise from et, ratio = (pi^2 + ln^2 et) /
(pi^2 + ln^2 en), Gn = cn sqrt(m_n Kn), Kt = Kn ratio m_t / m_n, Gt = ct
sqrt(m_t Kt).

Norms use a square root whose derivative at zero is taken as zero (bodies at
rest), as the GPU's dual-number square root does.

Parity: tests/test_torch_ref.py checks the forward values against the
oracle's own step and the gradient against central differences of the
oracle.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import oracle as O

PI2 = math.pi * math.pi
EPA_PLACEHOLDER = 1e20  # EPA face distances at or above this are placeholders (no depth resolved)


def _dot(a, b):
    return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]


def _cross(a, b):
    return torch.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                        a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                        a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], dim=-1)


def _sqrt0(x):
    """sqrt with derivative 0 where x == 0."""
    pos = x > 0
    return torch.where(pos, torch.sqrt(torch.where(pos, x, torch.ones_like(x))), torch.zeros_like(x))


def _qmul(a, b):
    """Hamilton product, (w, x, y, z) layout (gpu_Verlet.cu:43-54)."""
    v1, v2 = a[..., 1:], b[..., 1:]
    s = (v2 * a[..., :1] + v1 * b[..., :1]) + _cross(v1, v2)
    f = a[..., 0] * b[..., 0] - _dot(v1, v2)
    return torch.cat([f[..., None], s], dim=-1)


def _conj(q):
    return torch.cat([q[..., :1], -q[..., 1:]], dim=-1)


def _qrot(q1, v, q2):
    """Vector part of q1 (0, v) q2 (gpu_Verlet.cu:57-74)."""
    vq = torch.cat([torch.zeros_like(v[..., :1]), v], dim=-1)
    return _qmul(q1, _qmul(vq, q2))[..., 1:]


def _matv(M, v):
    """M [..., 3, 3] (row-major) times v [..., 3]."""
    return torch.stack([_dot(M[..., r, :], v) for r in range(3)], dim=-1)


def law_coeffs(Kn, en, et, mu):
    ln_en = torch.log(en)
    ln_et = torch.log(torch.as_tensor(et, dtype=torch.float64))
    A_en = PI2 + ln_en * ln_en
    return {"Kn": Kn, "mu": mu, "cn": (-2.0 * ln_en) / torch.sqrt(A_en),
            "ct": (-2.0 * ln_et) / torch.sqrt(PI2 + ln_et * ln_et), "ratio": (PI2 + ln_et * ln_et) / A_en}


def face_geometry(VA, VB):
    """VA, VB [K, 3, 3]: the face's support vertices on A and B (stored
    order). Returns depth [K], normal [K, 3], cpA [K, 3], cpB [K, 3]."""
    V = VA - VB
    Va, Vb, Vc = V[:, 0], V[:, 1], V[:, 2]
    nn = _cross(Vb - Va, Vc - Vb)
    n = nn / torch.sqrt(_dot(nn, nn))[:, None]
    depth = _dot(n, Va)
    proj = n * depth[:, None]
    v0, v1, v2 = Vb - Va, Vc - Va, proj - Va
    swap = _dot(_cross(v0, v1), n) < 0.0  # never with the stored order; kept as the oracle has it
    v0, v1 = torch.where(swap[:, None], v1, v0), torch.where(swap[:, None], v0, v1)
    v00, v01, v11 = _dot(v0, v0), _dot(v0, v1), _dot(v1, v1)
    v20, v21 = _dot(v2, v0), _dot(v2, v1)
    den = v00 * v11 - v01 * v01
    vv = (v11 * v20 - v01 * v21) / den
    ww = (v00 * v21 - v01 * v20) / den
    uu = 1.0 - ww - vv
    cpA = (VA[:, 0] * uu[:, None] + VA[:, 1] * vv[:, None]) + VA[:, 2] * ww[:, None]
    cpB = (VB[:, 0] * uu[:, None] + VB[:, 1] * vv[:, None]) + VB[:, 2] * ww[:, None]
    return depth, n, cpA, cpB


def contact_forces(st, sc, law_p, law_w, a, b, face):
    """Forces and torques of the colliding pairs (a[k], b[k]) with EPA faces
    face[k] (6 support indices), summed per row in pair order."""
    N = st["com"].shape[0]
    W = sc.n_walls
    F = torch.zeros((N, 3), dtype=torch.float64)
    T = torch.zeros((N, 3), dtype=torch.float64)
    if a.size == 0:
        return F, T
    at, bt = torch.from_numpy(a.astype(np.int64)), torch.from_numpy(b.astype(np.int64))
    fa = torch.from_numpy(face[:, :3].astype(np.int64))
    fb = torch.from_numpy(face[:, 3:].astype(np.int64))
    verts = st["verts"]
    VA = verts[at[:, None], fa]   # [K, 3, 3]
    VB = verts[bt[:, None], fb]
    depth, n, cpA, cpB = face_geometry(VA, VB)
    xA, xB = st["com"][at], st["com"][bt]
    vA, vB = st["v"][at], st["v"][bt]
    wA, wB = st["w"][at], st["w"][bt]
    mass = torch.from_numpy(sc.mass)
    isz = torch.from_numpy(sc.isize)
    wallA = torch.from_numpy(a < W)
    wallB = torch.from_numpy(b < W)
    imA = torch.where(wallA, torch.zeros(()), 1.0 / mass[at])
    imB = torch.where(wallB, torch.zeros(()), 1.0 / mass[bt])
    wall = wallA | wallB
    Kn = torch.where(wall, law_w["Kn"], law_p["Kn"])
    mu = torch.where(wall, law_w["mu"], law_p["mu"])
    cn = torch.where(wall, law_w["cn"], law_p["cn"])
    ct = torch.where(wall, law_w["ct"], law_p["ct"])
    ratio = torch.where(wall, law_w["ratio"], law_p["ratio"])
    cp = (cpA + cpB) / 2.0
    dA, dB = cp - xA, cp - xB
    dAn, dBn = _dot(dA, dA), _dot(dB, dB)
    imn = imA + imB
    mn = 1.0 / imn
    imt = (imn + dAn * torch.where(imA == 0.0, torch.zeros(()), 1.0 / isz[at])) + \
        dBn * torch.where(imB == 0.0, torch.zeros(()), 1.0 / isz[bt])
    mt = 1.0 / imt
    Gn = cn * torch.sqrt(mn * Kn)
    Kt = (Kn * ratio) * (mt / mn)
    Gt = ct * torch.sqrt(mt * Kt)
    flip = _dot(xB - xA, n) < 0.0
    n = torch.where(flip[:, None], -n, n)
    n = n / torch.sqrt(_dot(n, n))[:, None]
    disp_n = n * depth[:, None]
    u = ((vA + _cross(wA, cpA - xA)) - vB) - _cross(wB, cpB - xB)
    un = n * _dot(u, n)[:, None]
    ut = u - un
    utn = _sqrt0(_dot(ut, ut))
    t = torch.where((utn != 0.0)[:, None], ut / torch.where(utn != 0.0, utn, torch.ones_like(utn))[:, None],
                    torch.zeros_like(ut))
    Fn = (disp_n * -1.0) * Kn[:, None] - un * Gn[:, None]
    Fnn = _sqrt0(_dot(Fn, Fn))
    Ft = ut * -Gt[:, None]
    Ftn = _sqrt0(_dot(Ft, Ft))
    cap = Ftn > mu * Fnn
    Ft = torch.where(cap[:, None], ((t * Fnn[:, None]) * -1.0) * mu[:, None], Ft)
    Ff = Ft + Fn
    tA = _cross(cp - xA, Ff)
    tB = _cross(cp - xB, Ff * -1.0)
    # per-row sums in pair order: side A then side B of each contact
    rows = torch.stack([at, bt], dim=1).reshape(-1)
    F = F.index_add(0, rows, torch.stack([Ff, -Ff], dim=1).reshape(-1, 3))
    T = T.index_add(0, rows, torch.stack([tA, tB], dim=1).reshape(-1, 3))
    return F, T


def verlet(st, F, T, sc):
    """One Verlet step of the moving rows (gpu_Verlet.cu:164-403), vertices
    turned by the increment q q_old^-1. Returns the new state dict."""
    W, N, dt = sc.n_walls, sc.n_rows, sc.dt
    mv = slice(W, N)
    g = {k: v[mv] for k, v in st.items()}
    m = torch.from_numpy(sc.mass[mv])[:, None]
    I = torch.from_numpy(sc.I[mv])
    Iinv = torch.from_numpy(sc.Iinv[mv])
    F, T = F[mv], T[mv]
    v12, w12bf, wbf, w, com = g["v12"], g["w12bf"], g["wbf"], g["w"], g["com"]
    q, q12, dxall = g["q"], g["q12"], g["dxall"]
    v32 = v12 + (F / m) * dt
    tbf = _qrot(_conj(q), T, q)
    p2 = tbf - _cross(wbf, _matv(I, wbf))
    abf = _matv(Iinv, p2)
    w32bf = w12bf + abf * dt
    wn = _sqrt0(_dot(w, w))
    nz = (wn != 0.0)[:, None]
    wsafe = torch.where(wn != 0.0, wn, torch.ones_like(wn))[:, None]
    t11 = torch.where(nz, (w / wsafe) * torch.sin((wn[:, None] * dt) / 2.0), torch.zeros_like(w))
    t112 = torch.cos((wn * dt) / 2.0)
    q12 = _qmul(torch.cat([t112[:, None], t11], dim=-1), q12)
    v12 = v32
    w12bf = w32bf
    dx = v12 * dt
    dxall = dxall + dx
    w34 = _qrot(q12, w12bf + (abf * dt) / 4.0, _conj(q12))
    n34 = _sqrt0(_dot(w34, w34))
    nz34 = (n34 != 0.0)[:, None]
    s34 = torch.where(n34 != 0.0, n34, torch.ones_like(n34))[:, None]
    t1 = torch.where(nz34, (w34 / s34) * torch.sin((n34[:, None] * dt) / 4.0), torch.zeros_like(w34))
    t12 = torch.cos((n34 * dt) / 4.0)
    qold = q
    q = _qmul(torch.cat([t12[:, None], t1], dim=-1), q12)
    qall = _qmul(q, qold)
    wbf = w12bf + (abf * dt) / 2.0
    w = _qrot(q, wbf, _conj(q))
    vel = v12 + (F * dt) / m
    qr = _qmul(q, _conj(qold))
    box = sc.box
    c1 = com + dx
    out = ((c1[:, 0] > box[0]) | (c1[:, 0] < box[1]) | (c1[:, 1] > box[2]) | (c1[:, 1] < box[3]) |
           (c1[:, 2] > box[4]) | (c1[:, 2] < box[5]))
    nv = torch.from_numpy(sc.nverts[mv].astype(np.int64))
    V = g["verts"]
    vmax = V.shape[1]
    live = (torch.arange(vmax)[None, :] < nv[:, None])[..., None]
    qrv = qr[:, None, :].expand(-1, vmax, -1)
    r = _qrot(qrv, V - com[:, None, :], _conj(qrv))
    nverts_new = (r + com[:, None, :]) + dx[:, None, :]
    nverts_new = torch.where(out[:, None, None], nverts_new + dx[:, None, :] * -1.0, nverts_new)
    verts = torch.where(live, nverts_new, V)
    com = com + dx
    o3 = out[:, None]
    z = torch.zeros_like(vel)
    com = torch.where(o3, com + dx * -1.0, com)
    vel, v12 = torch.where(o3, z, vel), torch.where(o3, z, v12)
    w12bf, wbf, w = torch.where(o3, z, w12bf), torch.where(o3, z, wbf), torch.where(o3, z, w)
    new = {"com": com, "v": vel, "v12": v12, "w": w, "wbf": wbf, "w12bf": w12bf, "q": q, "q12": q12,
           "qall": qall, "dxall": dxall, "verts": verts}
    return {k: torch.cat([st[k][:W], new[k]], dim=0) for k in st}


def step(st, sc, law_p, law_w):
    """One step: topology from the oracle on the detached state, then the
    differentiable forces and Verlet. Returns (new state, n_contacts)."""
    npst = {k: np.ascontiguousarray(v.detach().numpy()) for k, v in st.items()}
    pairs = O.candidates(npst, sc)
    coll, depth, _, _, _, _, _, _ = O.contact_state(pairs, npst, sc)
    _, face = O.contact_faces(pairs, npst, sc)
    hit = (coll == 1) & (depth < EPA_PLACEHOLDER) & (depth > 0.0)
    a, b = pairs[hit, 0], pairs[hit, 1]
    F, T = contact_forces(st, sc, law_p, law_w, a, b, face[hit])
    F = F + torch.from_numpy(sc.fext)
    T = T + torch.from_numpy(sc.text)
    return verlet(st, F, T, sc), int(hit.sum())


def run(sc, init: dict, theta, steps: int):
    """Run `steps` steps from the numpy state `init` with theta = (Kn, en, mu)
    (a float64 tensor, possibly requiring grad) on both laws; returns the
    final state (torch tensors) and the contact count per step."""
    law_p = law_coeffs(theta[0], theta[1], sc.law_p.et, theta[2])
    law_w = law_coeffs(theta[0], theta[1], sc.law_w.et, theta[2])
    st = {k: torch.from_numpy(np.array(init[k], dtype=np.float64)) for k in O.STATE_KEYS}
    counts = []
    for _ in range(steps):
        st, k = step(st, sc, law_p, law_w)
        counts.append(k)
    return st, counts

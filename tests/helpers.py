"""Shared helpers of the test-suite (random states, comparisons)."""
from __future__ import annotations

import numpy as np


def random_motion(st: dict, sc, seed: int, vscale=50.0, wscale=200.0, quats=True):
    """Give the moving particles random velocities, angular velocities and
    (with quats) random orientation quaternions. Vertices are left as they are."""
    rng = np.random.Generator(np.random.PCG64(seed))
    W, N = sc.n_walls, sc.n_rows
    n = N - W
    st["v"][W:] = rng.normal(size=(n, 3)) * vscale
    st["v12"][W:] = st["v"][W:] + rng.normal(size=(n, 3))
    st["w"][W:] = rng.normal(size=(n, 3)) * wscale
    st["wbf"][W:] = rng.normal(size=(n, 3)) * wscale
    st["w12bf"][W:] = st["wbf"][W:] + rng.normal(size=(n, 3))
    if quats:
        for k in ("q", "q12", "qall"):
            q = rng.normal(size=(n, 4))
            q /= np.linalg.norm(q, axis=1, keepdims=True)
            st[k][W:] = q
    return st


def rel_err(a, b, floor=1e-300):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.maximum(np.abs(b), floor)
    return float(np.max(np.abs(a - b) / den)) if a.size else 0.0


def max_rel(a, b, scale=None):
    """max |a-b| / max(|b|, scale) with scale defaulting to max|b| * 1e-3."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.size == 0:
        return 0.0
    s = scale if scale is not None else max(float(np.max(np.abs(b))) * 1e-3, 1e-300)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), s)))


def gpu_contact_pairs(verts, nverts, com, pairs, law_p=(1e4, 0.5, 0.5, 0.3, 0), law_w=None, n_walls=0, dt=1e-6,
                      vel=None, angvel=None, q=None, isize=None, entry="mwd_contact_pairs", n_rows=None):
    """mwd_contact_pairs (the per-pair entry point, through the C ABI) on a
    pair list indexed into per-body arrays: gathers each pair's rows into the
    reference's per-pair layout and returns the per-pair outputs and the
    per-body force/torque sums as numpy arrays."""
    import ctypes as C

    import torch

    from mwdiem_amd import _lib
    dev = torch.device("cuda", 0)
    pairs = np.asarray(pairs, np.int64).reshape(-1, 2)
    P, vmax = pairs.shape[0], verts.shape[1]
    N = verts.shape[0] if n_rows is None else n_rows
    a, b = pairs[:, 0], pairs[:, 1]
    z3 = np.zeros((N, 3))
    vel = z3 if vel is None else vel
    angvel = z3 if angvel is None else angvel
    q = np.tile([1.0, 0, 0, 0], (N, 1)) if q is None else q
    isize = np.ones(N) if isize is None else isize

    def t(x, dt_=torch.float64):
        return torch.from_numpy(np.ascontiguousarray(x)).to(dev, dt_)

    args = [t(com[a]), t(com[b]), t(verts[a]), t(verts[b]), t(vel[a]), t(vel[b]), t(angvel[a]), t(angvel[b]),
            t(nverts[a], torch.int32), t(nverts[b], torch.int32), t(a, torch.int32), t(b, torch.int32),
            t(np.zeros((P, 4))), t(q[a]), t(isize[a]), t(isize[b])]
    out = dict(collision=torch.full((P,), -7, dtype=torch.int32, device=dev),
               epa_status=torch.zeros(P, dtype=torch.int32, device=dev),
               depth=torch.zeros(P, dtype=torch.float64, device=dev),
               normal=torch.zeros((P, 3), dtype=torch.float64, device=dev),
               cpA=torch.zeros((P, 3), dtype=torch.float64, device=dev),
               cpB=torch.zeros((P, 3), dtype=torch.float64, device=dev),
               disp=torch.zeros((P, 4), dtype=torch.float64, device=dev))
    F = torch.zeros((N, 3), dtype=torch.float64, device=dev)
    T = torch.zeros((N, 3), dtype=torch.float64, device=dev)
    law_w = law_p if law_w is None else law_w
    lp = _lib.law(*law_p[:4], dt, int(law_p[4]))
    lw = _lib.law(*law_w[:4], dt, int(law_w[4]))
    rc = getattr(_lib.load(), entry)(P, vmax, *[x.data_ptr() for x in args], C.byref(lp), C.byref(lw), n_walls,
                                       out["collision"].data_ptr(), out["epa_status"].data_ptr(),
                                       out["depth"].data_ptr(), out["normal"].data_ptr(), out["cpA"].data_ptr(),
                                       out["cpB"].data_ptr(), out["disp"].data_ptr(), F.data_ptr(), T.data_ptr(),
                                       N, None)
    _lib.check(rc, entry)
    torch.cuda.synchronize()
    res = {k: v.cpu().numpy() for k, v in out.items()}
    res["force"], res["torque"] = F.cpu().numpy(), T.cpu().numpy()
    return res

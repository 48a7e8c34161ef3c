"""Writes tests/golden/rare_paths.npz (run from the repo root):

    python tests/golden/make_rare_paths.py

Hull pairs that drive GJK/EPA down the reference's rare branches, found by a
seeded search on the oracle, with the oracle's results:
  * flat     -- CorrectFlat fires (gpu_fine_contact_detection.cu:256-268,
                1119-1128): hulls about 1e-3 mm across, whose simplex
                determinants fall under the 1e-6 threshold;
  * epa_cap  -- EPA stops at its expansion cap (.cu:443-447, status bit 1):
                deep overlaps of 20-32-vertex hulls;
  * gjk_cap  -- the tetrahedron loop runs out its 50 iterations and the pair
                keeps collision -1 (.cu:1104,1287): hulls that touch, B
                placed on A's support plane to within 1e-9 relative.
Every pair is drawn from its own generator PCG64([family seed, trial]), so a
case is reproduced from its trial number alone.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
from scipy.spatial import ConvexHull

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import oracle as O  # noqa: E402

OUT = Path(__file__).resolve().parent / "rare_paths.npz"
VMAX = 32


def hull(rng, n, axes):
    p = rng.normal(size=(n, 3))
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    p *= axes
    return p[ConvexHull(p).vertices]


def draw(family: str, trial: int, with_dir: bool = False):
    """One hull pair of a family: (A, B) vertex arrays (and, with with_dir,
    the direction along which B was placed on A's support plane; zero for
    the other families)."""
    seed = {"flat": 101, "epa_cap": 102, "gjk_cap": 103}[family]
    rng = np.random.Generator(np.random.PCG64([seed, trial]))
    d = np.zeros(3)
    if family == "flat":
        s = 10 ** rng.uniform(-3.2, -2.6)
        A = hull(rng, int(rng.integers(6, 21)), rng.uniform(0.6, 1.2, 3) * s)
        B = hull(rng, int(rng.integers(6, 21)), rng.uniform(0.6, 1.2, 3) * s)
        e = rng.normal(size=3)
        B = B + e / np.linalg.norm(e) * rng.uniform(0.0, 1.2) * s
    elif family == "epa_cap":
        A = hull(rng, int(rng.integers(20, 33)), rng.uniform(0.6, 1.2, 3))
        B = hull(rng, int(rng.integers(20, 33)), rng.uniform(0.6, 1.2, 3)) + rng.normal(size=3) * 0.05
    else:
        A = hull(rng, int(rng.integers(4, 33)), rng.uniform(0.3, 1.5, 3))
        B = hull(rng, int(rng.integers(4, 33)), rng.uniform(0.3, 1.5, 3))
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        gap = (A @ d).max() - (B @ d).min()
        B = B + d * gap * (1.0 + rng.choice([-1.0, 1.0]) * 10 ** rng.uniform(-16, -9))
    return (A, B, d) if with_dir else (A, B)


def pack(pairs_AB):
    P = len(pairs_AB)
    verts = np.zeros((2 * P, VMAX, 3))
    nv = np.zeros(2 * P, np.int32)
    com = np.zeros((2 * P, 3))
    for i, (A, B) in enumerate(pairs_AB):
        for r, V in ((2 * i, A), (2 * i + 1, B)):
            verts[r, :len(V)] = V
            nv[r] = len(V)
            com[r] = V.mean(axis=0)
    pairs = np.stack([np.arange(0, 2 * P, 2), np.arange(1, 2 * P, 2)], axis=1).astype(np.int32)
    return verts, nv, com, pairs


def search(family: str, want: int, trials: int, pred):
    found = []
    batch = 2000
    for t0 in range(0, trials, batch):
        ts = list(range(t0, t0 + batch))
        verts, nv, com, pairs = pack([draw(family, t) for t in ts])
        c, it, fl, es, ei = O.gjk_trace(pairs, com, verts, nv)
        for k, t in enumerate(ts):
            if pred(c[k], it[k], fl[k], es[k]):
                found.append(t)
        if len(found) >= want:
            break
    return found[:want]


def main():
    fams = {
        "flat": search("flat", 24, 20000, lambda c, it, fl, es: fl == 1 and c == 1 and es == 0),
        "epa_cap": search("epa_cap", 8, 40000, lambda c, it, fl, es: c == 1 and (es & 1) and not (es & 2)),
        "gjk_cap": search("gjk_cap", 3, 400000, lambda c, it, fl, es: c == -1),
    }
    cases = [(f, t) for f, ts in fams.items() for t in ts]
    verts, nv, com, pairs = pack([draw(f, t) for f, t in cases])
    c, it, fl, es, ei = O.gjk_trace(pairs, com, verts, nv)
    out = {}
    for i, (A, B) in enumerate([draw(f, t) for f, t in cases]):
        cc, dep, n, ca, cb, st = O.gjk_epa(A, B, A.mean(axis=0), B.mean(axis=0))
        out.setdefault("depth", []).append(dep)
        out.setdefault("normal", []).append(n)
        out.setdefault("cpA", []).append(ca)
        out.setdefault("cpB", []).append(cb)
    dirs = np.array([draw(f, t, with_dir=True)[2] for f, t in cases])
    np.savez_compressed(OUT, family=np.array([f for f, _ in cases]), trial=np.array([t for _, t in cases]),
                        touch_dir=dirs,
                        verts=verts, nverts=nv, com=com, pairs=pairs, collision=c, gjk_iters=it, flat=fl,
                        epa_status=es, epa_iters=ei, depth=np.array(out["depth"]), normal=np.array(out["normal"]),
                        cpA=np.array(out["cpA"]), cpB=np.array(out["cpB"]))
    for f, ts in fams.items():
        print(f, len(ts), ts)


if __name__ == "__main__":
    main()

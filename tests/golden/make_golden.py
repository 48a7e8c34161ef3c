"""Writes the golden fixtures of the test-suite (run from the repo root):

    python tests/golden/make_golden.py

Each fixture holds a small seeded scene (inputs, as plain arrays) and the
oracle's outputs on it: the candidate list, collision codes, EPA results and
the state after a few full steps. The reference's contact and Verlet kernels
cannot run here (they need pycuda, nvcc and a CUDA device; its pure-Python
modules -- the cell hash, the writers, the particle class -- do import), so
these vectors come from the oracle, which the pins and the reference-derived
fixtures check; they hold the oracle and the GPU path against silent
changes.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from mwdiem_amd import scene  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = Path(__file__).resolve().parent


def fixture(sc, name, steps):
    st = O.initial_state(sc)
    W = sc.n_walls
    rng = np.random.Generator(np.random.PCG64(77))
    st["v"][W:] = rng.normal(size=(sc.n_particles, 3)) * 200.0
    st["v12"][W:] = st["v"][W:]
    st0 = {k: v.copy() for k, v in st.items()}
    pairs = O.candidates(st, sc)
    coll, depth, nrm, ca, cb, F, T, est = O.contact_state(pairs, st, sc)
    for _ in range(steps):
        O.step(st, sc)
    d = sc.to_arrays()
    d.update({f"in_{k}": v for k, v in st0.items()})
    d.update(pairs=pairs, collision=coll, depth=depth, normal=nrm, cpA=ca, cpB=cb, force=F, torque=T,
             steps=np.array(steps))
    d.update({f"out_{k}": st[k] for k in ("com", "v", "w", "q", "q12", "qall", "verts")})
    np.savez_compressed(OUT / f"{name}.npz", **d)
    print(name, "pairs", len(pairs), "contacts", int((coll == 1).sum()))


def main():
    base = scene.make_scene(24, (6.0, 6.0, 12.0), 41, name="golden")
    # the reference law amplifies a one-ulp difference about tenfold per step
    # on this overlapped bed (its damping uses mu as a mass), so its fixture
    # stops after 2 steps; the physical law is stable over 10
    fixture(scene.compact_dense(base, gap=-0.55), "dense_ref", 2)
    phys = scene.make_scene(24, (6.0, 6.0, 12.0), 43, law=scene.Law4(mode=scene.PHYSICAL), name="golden-phys")
    fixture(scene.compact_dense(phys, gap=-0.55), "dense_phys", 10)


if __name__ == "__main__":
    main()

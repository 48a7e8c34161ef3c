"""Writes tests/golden/epa_unresolved.npz (run from the repo root; about six
minutes on one core):

    python tests/golden/make_epa_unresolved.py

The oracle settles the 4,096-hull pile (seed 5, no pour velocity, physical
law) step by step. At step 29,935 the pair (497, 47) is the first whose EPA
ends on a placeholder face (distance 3e20 of a deleted face). The fixture
keeps the two vertex sets and centres of that step: GJK reports contact, EPA
returns the placeholder depth and sets status bit 4.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from mwdiem_amd import scene  # noqa: E402
from oracle import oracle as O  # noqa: E402

STEPS = 29_934


def main():
    sc, _ = scene.config_pile(n=4096, v_pour=0.0)
    st = O.initial_state(sc)
    for _ in range(STEPS):
        O.step(st, sc)
    a, b = 497, 47
    VA, VB = st["verts"][a, :sc.nverts[a]], st["verts"][b, :sc.nverts[b]]
    c, depth, n, ca, cb, status = O.gjk_epa(VA, VB, st["com"][a], st["com"][b])
    assert c == 1 and depth >= 1e20 and status & 4, (c, depth, status)
    np.savez_compressed(Path(__file__).resolve().parent / "epa_unresolved.npz", VA=VA, VB=VB, comA=st["com"][a],
                        comB=st["com"][b], collision=np.array(c), depth=np.array(depth), status=np.array(status))


if __name__ == "__main__":
    main()

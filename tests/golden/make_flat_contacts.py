"""Writes tests/golden/flat_contacts.npz: hull pairs in shallow contact on
which the reference's GJK/EPA returns a depth of about a hull's size.

The pairs were found in a pile of 4,096 hulls poured with the physical law
(mwdiem_amd.scene.config_pour(n=4096), 150,000 steps, then stepped on): at
these steps one contact's force jumped by four orders of magnitude. Each
input file is the full state (Stepper.numpy_state()) of the step before such
a jump; `--pair` names the two rows of each file's contact. The fixture keeps
the box's walls and the listed hulls only (a scene of 6 + 2k rows with the
pour scene's grid and laws), each hull with its state row from its file.

    python tests/golden/make_flat_contacts.py state0.npz:2356,1645 \
        state1.npz:1289,1567 state2.npz:3205,2091 state3.npz:2104,1721
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[2]))

from mwdiem_amd import scene  # noqa: E402
from mwdiem_amd.stepper import STATE_KEYS  # noqa: E402

OUT = Path(__file__).resolve().parent / "flat_contacts.npz"


def main(args):
    sc, _, _ = scene.config_pour(n=4096)
    W = sc.n_walls
    src = []
    for a in args:
        path, rows = a.rsplit(":", 1)
        ra, rb = (int(x) for x in rows.split(","))
        src.append((np.load(path), ra, rb))
    rows = list(range(W)) + [r for _, ra, rb in src for r in (ra, rb)]
    d = sc.to_arrays()
    per_row = ("com", "verts", "nverts", "mass", "I", "Iinv", "isize", "fext", "text", "wall_n", "wall_c")
    out = {k: (np.asarray(v)[rows] if k in per_row else v) for k, v in d.items()}
    state = {k: np.asarray(src[0][0][k])[:W].copy() for k in STATE_KEYS}
    for st, ra, rb in src:
        for k in STATE_KEYS:
            state[k] = np.concatenate([state[k], np.asarray(st[k])[[ra, rb]]])
    # the scene's centres are the state's (the stepper reads both)
    out["com"] = state["com"].copy()
    out["verts"] = state["verts"].copy()
    for k in STATE_KEYS:
        out["state_" + k] = state[k]
    n = len(src)
    out["pairs"] = np.array([[W + 2 * i, W + 2 * i + 1] for i in range(n)], np.int32)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({n} pairs, {len(rows)} rows)")


if __name__ == "__main__":
    main(sys.argv[1:])

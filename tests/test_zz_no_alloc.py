"""mwd_contact_pairs_noalloc (the allocation-free form of mwd_contact_pairs,
its temporaries in a fixed device array) against mwd_contact_pairs through
the C ABI: identical per-pair outputs and force/torque sums, bit for bit, on
two calls in a row -- the dense bed's candidate pairs, then a smaller input
(the first third of them) whose call finds the previous call's contents in
the array -- and oversized inputs refused without a launch."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _inputs(sc, seed):
    from tests.helpers import random_motion
    st = random_motion(O.initial_state(sc), sc, seed)
    pairs = O.candidates(st, sc)
    return st, pairs


def _run(entry, sc, st, pairs):
    from tests.helpers import gpu_contact_pairs
    law = (sc.law_p.Kn, sc.law_p.en, sc.law_p.et, sc.law_p.mu, 0)
    return gpu_contact_pairs(st["verts"], sc.nverts, st["com"], pairs, law_p=law, n_walls=sc.n_walls, dt=sc.dt,
                             vel=st["v"], angvel=st["w"], q=st["q"], isize=sc.isize, entry=entry)


def test_noalloc_matches_contact_pairs_on_two_calls(gpu, dense_scene):
    sc = dense_scene
    st, pairs = _inputs(sc, 3)
    small = pairs[: max(1, pairs.shape[0] // 3)]
    for name, pp in (("full", pairs), ("smaller", small)):
        print(f"no-alloc: {name} call, {pp.shape[0]} pairs", flush=True)
        ref = _run("mwd_contact_pairs", sc, st, pp)
        got = _run("mwd_contact_pairs_noalloc", sc, st, pp)
        assert int((ref["collision"] == 1).sum()) > 0
        for k in ref:
            assert np.array_equal(got[k], ref[k]), (name, k)


def test_noalloc_refuses_inputs_beyond_its_limits(gpu, dense_scene):
    from mwdiem_amd import _lib
    sc = dense_scene
    st, pairs = _inputs(sc, 4)
    big = np.tile(pairs, (_lib.NOALLOC_MAX_PAIRS // pairs.shape[0] + 1, 1))
    assert big.shape[0] > _lib.NOALLOC_MAX_PAIRS
    with pytest.raises(_lib.MwdError) as e:
        _run("mwd_contact_pairs_noalloc", sc, st, big)
    assert e.value.code == _lib.MWD_E_INVALID
    from tests.helpers import gpu_contact_pairs
    law = (sc.law_p.Kn, sc.law_p.en, sc.law_p.et, sc.law_p.mu, 0)
    with pytest.raises(_lib.MwdError):
        gpu_contact_pairs(st["verts"], sc.nverts, st["com"], pairs, law_p=law, n_walls=sc.n_walls,
                          entry="mwd_contact_pairs_noalloc", n_rows=_lib.NOALLOC_MAX_PARTICLES + 1)
    # and the entry point still works after the refusals
    ref = _run("mwd_contact_pairs", sc, st, pairs)
    got = _run("mwd_contact_pairs_noalloc", sc, st, pairs)
    assert np.array_equal(got["force"], ref["force"]) and np.array_equal(got["collision"], ref["collision"])

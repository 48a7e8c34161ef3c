"""Drop-in for LIBRARIES/LIBRARY_Coarse_contact_detection.py: Linked_cell_hash
with the reference signature (:84) and behaviour -- it appends the candidate
pairs [A, B] of particles num_part_min..num_part_max-1 to the caller's list
and returns that list -- computed on the GPU (cell keys counted into a
hash table of occupied cells, per-cell member lists, pair emission) in
exactly the reference's order: cells in order of their lowest id, members in
CPython set order, same-cell pairs then the 7 forward offsets.

Like the reference it reads only hashh_arr[i] and assumes
particle[i].particle_id == i.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import _lib
from ..stepper import _copy_from_device
from ._common import device


def Linked_cell_hash(particle, numcells_x, numcells_y, num_part_min, num_part_max, pairstocheck, hashh_arr):
    dev = device()
    pmin, pmax = int(num_part_min), int(num_part_max)
    keys_np = np.asarray(hashh_arr).reshape(-1).astype(np.int64)
    if pmax <= pmin:
        return pairstocheck
    n_rows = max(pmax, keys_np.shape[0])
    k = keys_np[pmin:pmax]
    kmin, kmax = int(k.min()), int(k.max())
    ncx, ncy = int(numcells_x), int(numcells_y)
    # neighbour lookups reach kmax + ncx*ncy + ncx + 1; the span must hold
    # every key that exists, targets outside it simply do not exist
    sc = _lib.Scene()
    sc.row0, sc.n_walls, sc.n_active = 0, pmin, pmax
    for i in range(6):
        sc.box[i] = 0.0
    for i in range(3):
        sc.cell[i] = 1.0
    sc.ncx, sc.ncy, sc.kmin, sc.kspan = ncx, ncy, kmin, kmax - kmin + 1
    sc.wall_thr = -1.0
    sc.law_p = _lib.law(1.0, 0.5, 0.5, 0.5, 1.0)
    sc.law_w = sc.law_p
    n = pmax - pmin
    pair_cap = max(1024, 64 * n)
    lib = _lib.load()
    keys = torch.from_numpy(np.ascontiguousarray(np.pad(keys_np, (0, n_rows - keys_np.shape[0])))).to(dev)
    com = torch.zeros((n_rows, 3), dtype=torch.float64, device=dev)
    st = _lib.State()
    st.n_rows, st.vmax = n_rows, 1
    st.com = com.data_ptr()
    st.wall_n = None
    st.wall_c = None
    stream = torch.cuda.current_stream().cuda_stream or None
    while True:
        ws = C.c_void_p()
        _lib.check(lib.mwd_workspace_create(C.byref(ws), dev.index, C.byref(sc), 1, n_rows, 1, pair_cap, 1, 1.0,
                                            stream), "mwd_workspace_create")
        try:
            _lib.check(lib.mwd_candidates(ws, C.byref(st), keys.data_ptr()), "mwd_candidates")
            status = _lib.Status()
            _lib.check(lib.mwd_step_status(ws, C.byref(status)), "mwd_step_status")
            if status.flags & _lib.F_CELL:
                raise _lib.MwdError(_lib.MWD_E_CELL, "Linked_cell_hash",
                                    f"a cell holds more than {_lib.MAX_CELL} particles")
            if status.n_pairs > pair_cap:
                pair_cap = int(status.n_pairs)
                continue
            p = C.c_void_p()
            _lib.check(lib.mwd_pairs_view(ws, C.byref(p), None), "mwd_pairs_view")
            pairs = _copy_from_device(p.value, 2 * status.n_pairs, np.int32).reshape(-1, 2)
        finally:
            lib.mwd_workspace_destroy(ws)
        break
    pairstocheck.extend(pairs.tolist())
    return pairstocheck

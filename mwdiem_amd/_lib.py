"""ctypes binding of libmwdiem_hip.so (the C ABI in include/mwdiem_hip.h).

The library is loaded from this package directory. There is no fallback: if
the library is missing or a call fails, a MwdError is raised.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libmwdiem_hip.so"

MWD_OK = 0
MWD_E_INVALID = -1
MWD_E_HIP = -2
MWD_E_CAPACITY = -3
MWD_E_KEYRANGE = -4
MWD_E_CELL = -5

F_PAIR_OVERFLOW = 1
F_CONTACT_OVERFLOW = 2
F_KEYRANGE = 4
F_CELL = 8
F_EPA_CAP = 16
F_EPA_FACES = 32
F_ROW_CONTACTS = 64
F_EPA_UNRESOLVED = 128
F_TAPE = 256
FATAL_FLAGS = F_PAIR_OVERFLOW | F_CONTACT_OVERFLOW | F_KEYRANGE | F_CELL | F_ROW_CONTACTS

ROW_CONTACTS = 32

EPA_MAX_FACES = 128
MAX_CELL = 76

NOALLOC_MAX_PAIRS = 65536
NOALLOC_MAX_PARTICLES = 262144


class MwdError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str = ""):
        super().__init__(f"{where} failed with code {code}: {msg}")
        self.code = code


LAW_REF = 0
LAW_PHYSICAL = 1


class Law(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("Kn", "mu", "ln_en", "ln_et", "A_en", "tc", "Gn", "Kt_num", "inv_mu", "dt",
                                          "cn", "ct", "ratio")] + [("mode", C.c_int32), ("pad_", C.c_int32)]


N_TAN = 3
TAN_ARRAYS = ("com", "v", "v12", "w", "wbf", "w12bf", "q", "q12", "qall", "dxall", "verts")


class LawTan(C.Structure):
    _fields_ = [(n, C.c_double * N_TAN) for n in ("Kn", "mu", "Gn", "Kt_num", "inv_mu", "ln_et", "cn", "ct",
                                                   "ratio")]


class Scene(C.Structure):
    _fields_ = [
        ("row0", C.c_int64), ("n_walls", C.c_int64), ("n_active", C.c_int64),
        ("box", C.c_double * 6), ("cell", C.c_double * 3),
        ("ncx", C.c_int64), ("ncy", C.c_int64), ("kmin", C.c_int64), ("kspan", C.c_int64),
        ("wall_thr", C.c_double), ("law_p", Law), ("law_w", Law),
    ]


_P = C.c_void_p


class State(C.Structure):
    _fields_ = [
        ("n_rows", C.c_int64), ("vmax", C.c_int32),
        ("com", _P), ("v", _P), ("v12", _P), ("w", _P), ("wbf", _P), ("w12bf", _P),
        ("q", _P), ("q12", _P), ("qall", _P), ("dxall", _P), ("verts", _P), ("nverts", _P),
        ("mass", _P), ("I", _P), ("Iinv", _P), ("isize", _P), ("fext", _P), ("text", _P),
        ("wall_n", _P), ("wall_c", _P), ("radius", _P),
    ]


class Status(C.Structure):
    _fields_ = [("n_pairs", C.c_int64), ("n_wall_pairs", C.c_int64), ("n_contacts", C.c_int64),
                ("n_cells", C.c_int64), ("flags", C.c_uint32), ("n_sphere_pass", C.c_int64)]


class Tuning(C.Structure):
    _fields_ = [("cell_pass", C.c_int32), ("gjk_group", C.c_int32), ("gjk_blocks", C.c_int32),
                ("sphere_block", C.c_int32), ("verlet_block", C.c_int32), ("cell_items", C.c_int32),
                ("verlet_waves", C.c_int32), ("vertex_pack", C.c_int32), ("epa_blocks", C.c_int32),
                ("epa_fused", C.c_int32), ("fill_cells", C.c_int32),
                ("gjk_buckets", C.c_int32), ("fill_split", C.c_int32)]


HISTORY_RESET = 0
HISTORY_PERSISTENT = 1
REDUCE_FP64 = 0
REDUCE_FP32 = 1
CELLS_TABLE = 0
CELLS_SORTED = 1


# name -> (restype, argtypes)
_SIGS = {
    "mwd_last_error": (C.c_char_p, []),
    "mwd_version": (C.c_int, []),
    "mwd_law_init": (C.c_int, [C.POINTER(Law), C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]),
    "mwd_law_init_mode": (C.c_int, [C.POINTER(Law), C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                    C.c_int32]),
    "mwd_workspace_create": (C.c_int, [C.POINTER(_P), C.c_int, C.POINTER(Scene), C.c_int32, C.c_int64, C.c_int32,
                                       C.c_int64, C.c_int64, C.c_double, _P]),
    "mwd_workspace_destroy": (C.c_int, [_P]),
    "mwd_set_wall_motion": (C.c_int, [_P, C.c_int32]),
    "mwd_set_spawn": (C.c_int, [_P, C.c_int64, C.c_int64, C.c_int64]),
    "mwd_set_reduction": (C.c_int, [_P, C.c_int32]),
    "mwd_set_tangential_history": (C.c_int, [_P, C.c_int32]),
    "mwd_get_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "mwd_set_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "mwd_verlet": (C.c_int, [_P] * 16 + [C.c_int32, _P, C.c_double, _P, C.c_int64, C.c_int64, _P]),
    "mwd_contact_pairs": (C.c_int, [C.c_int64, C.c_int32] + [_P] * 16 + [_P, _P, C.c_int64] + [_P] * 9
                          + [C.c_int64, _P]),
    "mwd_contact_pairs_noalloc": (C.c_int, [C.c_int64, C.c_int32] + [_P] * 16 + [_P, _P, C.c_int64] + [_P] * 9
                                  + [C.c_int64, _P]),
    "mwd_cell_keys": (C.c_int, [_P, C.c_int64, _P, C.c_int64, C.c_int64, _P, _P]),
    "mwd_candidates": (C.c_int, [_P, C.POINTER(State), _P]),
    "mwd_pairs_view": (C.c_int, [_P, C.POINTER(_P), C.POINTER(_P)]),
    "mwd_fine_contact": (C.c_int, [_P, C.POINTER(State)]),
    "mwd_contacts_view": (C.c_int, [_P] + [C.POINTER(_P)] * 11),
    "mwd_reduce_verlet": (C.c_int, [_P, C.POINTER(State)]),
    "mwd_step": (C.c_int, [_P, C.POINTER(State)]),
    "mwd_step_status": (C.c_int, [_P, C.POINTER(Status)]),
    "mwd_clear_flags": (C.c_int, [_P]),
    "mwd_set_timing": (C.c_int, [_P, C.c_int]),
    "mwd_radius_init": (C.c_int, [_P, _P, _P, C.c_int64, C.c_int32, _P, _P]),
    "mwd_hull_bodies": (C.c_int, [_P, _P, C.c_int64, C.c_int32, C.c_double, _P, _P, _P, _P, _P, _P, _P, _P]),
    "mwd_copy_arrays": (C.c_int, [C.c_int32, C.POINTER(_P), C.POINTER(_P), C.POINTER(C.c_int64), _P]),
    "mwd_stage_times": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "mwd_kernel_times": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "mwd_law_tangent": (C.c_int, [C.POINTER(LawTan), C.c_double, C.c_double, C.c_double, C.c_double,
                                  C.POINTER(C.c_double)]),
    "mwd_set_law_tangents": (C.c_int, [_P, C.POINTER(LawTan)]),
    "mwd_step_tangent": (C.c_int, [_P, C.POINTER(State), C.POINTER(_P)]),
    "mwd_step_adjoint": (C.c_int, [_P, C.POINTER(State), C.POINTER(_P), _P]),
    "mwd_tape_init": (C.c_int, [_P, C.c_int64, C.POINTER(C.c_int64)]),
    "mwd_tape_rewind": (C.c_int, [_P]),
    "mwd_step_record": (C.c_int, [_P, C.POINTER(State)]),
    "mwd_step_adjoint_taped": (C.c_int, [_P, C.POINTER(State), C.POINTER(_P), _P]),
}

N_KTIMED = 5
KTIMED_NAMES = ("k_cell_fill", "k_sphere", "k_gjk", "k_epa_force", "k_reduce_verlet_v")
KTIMED_STAGE = ("emit", "sphere", "gjk", "epa+law", "reduce+verlet")

N_STAGES = 10
STAGE_NAMES = ("walls", "keys+sort", "cells", "emit", "pair-count", "sphere", "gjk", "compact", "epa+law",
               "reduce+verlet")

EXPORTED = tuple(_SIGS)

_lib = None


def load():
    """Load the library (once). Raises MwdError/OSError when it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    # torch first: its HIP runtime (soname libamdhip64.so.7) then satisfies
    # the library's dependency, so one runtime serves tensors, streams and kernels
    import torch  # noqa: F401
    if not LIB_PATH.exists():
        raise OSError(f"{LIB_PATH} is not built; run __graft_entry__.build() (or make -C mwdiem_amd/csrc)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, where: str) -> None:
    if rc != MWD_OK:
        msg = load().mwd_last_error()
        raise MwdError(rc, where, msg.decode() if msg else "")


def law(Kn: float, en: float, et: float, mu: float, dt: float, mode: int = LAW_REF) -> Law:
    L = Law()
    check(load().mwd_law_init_mode(C.byref(L), Kn, en, et, mu, dt, int(mode)), "mwd_law_init_mode")
    return L


def law_tangent(Kn: float, en: float, et: float, mu: float, seed) -> LawTan:
    """Derivatives of a law's coefficients; seed[i][k] = d(Kn, en, et, mu)[i] / d theta_k."""
    import numpy as np
    sd = np.ascontiguousarray(np.asarray(seed, dtype=np.float64).reshape(4, N_TAN))
    T = LawTan()
    check(load().mwd_law_tangent(C.byref(T), Kn, en, et, mu, sd.ctypes.data_as(C.POINTER(C.c_double))),
          "mwd_law_tangent")
    return T


def ptr(t) -> int | None:
    """Device address of a torch tensor (None for None)."""
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return t.data_ptr()


def stream_handle(torch_stream=None) -> int | None:
    import torch
    s = torch_stream if torch_stream is not None else torch.cuda.current_stream()
    return s.cuda_stream or None


MAX_COPY = 16


def copy_arrays(dst, src, stream=None) -> None:
    """dst[i].copy_(src[i]) for lists of contiguous device tensors of equal
    sizes (at most MAX_COPY), in one kernel launch on `stream` (a HIP stream
    handle; the current stream when None)."""
    n = len(dst)
    if n != len(src) or n > MAX_COPY:
        raise ValueError(f"copy_arrays: {n} destinations for {len(src)} sources (at most {MAX_COPY})")
    for d, s in zip(dst, src):
        if d.nbytes != s.nbytes or not d.is_contiguous() or not s.is_contiguous():
            raise ValueError("copy_arrays: arrays must be contiguous and of equal byte sizes")
    D = (_P * n)(*[d.data_ptr() for d in dst])
    S = (_P * n)(*[s.data_ptr() for s in src])
    B = (C.c_int64 * n)(*[d.nbytes for d in dst])
    check(load().mwd_copy_arrays(n, D, S, B, stream if stream is not None else stream_handle()), "mwd_copy_arrays")


def loaded_path() -> str:
    return os.fspath(LIB_PATH)

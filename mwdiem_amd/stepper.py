"""Device-resident time stepping: the reference's time-loop body
(MWDiem.py:805-1001) as one stream-ordered sequence of HIP kernels.

One Stepper holds one or many independent scenes (the batched form): rows of
all scenes are concatenated, each scene keeps its own walls, grid, law and
box, and every stage runs as one launch over all scenes. State lives in
PyTorch-ROCm tensors (FP64, the reference's array layouts); the kernels are
reached through the C ABI (mwdiem_amd/_lib.py).
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np
import torch

from . import _lib
from .scene import SceneData

STATE_KEYS = ("com", "v", "v12", "w", "wbf", "w12bf", "q", "q12", "qall", "dxall", "verts")
CONST_KEYS = ("nverts", "mass", "I", "Iinv", "isize", "fext", "text", "wall_n", "wall_c")


def _require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("mwdiem_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    _lib.load()


class Stepper:
    """Device-resident DEM step over one or more scenes.

    pair_cap / contact_cap bound the candidate and contact buffers; a step
    that exceeds them raises (after the step) from check(). spawn (a
    scene.Spawn, one scene only): the scene's rows join the run on that
    schedule, counted on the device from the first step.
    """

    def __init__(self, scenes: SceneData | Sequence[SceneData], device: int = 0, pair_cap: int | None = None,
                 contact_cap: int | None = None, stream: torch.cuda.Stream | None = None, sphere_test: bool = True,
                 reduce_fp32: bool = False, persistent_history: bool = False, spawn=None):
        _require_gpu()
        if isinstance(scenes, SceneData):
            scenes = [scenes]
        self.scenes = list(scenes)
        if len({sc.dt for sc in self.scenes}) != 1:
            raise ValueError("all scenes of a Stepper share one dt")
        self.dt = self.scenes[0].dt
        self.device = torch.device("cuda", device)
        # a stream of its own (default priority): steps can be captured in a graph
        self.stream = stream or torch.cuda.Stream(device=self.device)
        self.vmax = max(sc.vmax for sc in self.scenes)
        rows = [sc.n_rows for sc in self.scenes]
        self.row0 = np.concatenate([[0], np.cumsum(rows)[:-1]]).astype(np.int64)
        self.n_rows = int(sum(rows))
        n_mov = sum(sc.n_particles for sc in self.scenes)
        n_wall_tests = sum(sc.n_walls * sc.n_particles for sc in self.scenes)
        # about 32 candidates per particle on a grid of cells one particle
        # extent wide with the reference's 7 neighbour offsets (48 leaves some
        # room); the physical law's 13 offsets visit 13/7 as many neighbour
        # cells, and a wider grid holds more per cell. A scene may say what
        # it needs (scene.pairs_per_particle, e.g. a packed pile).
        cand = sum(sc.n_particles * (sc.pairs_per_particle if sc.pairs_per_particle else
                                     48 * (13.0 / 7.0 if sc.law_p.mode != 0 else 1.0)
                                     * max(1.0, float(np.prod(sc.cell)) / sc.max_particlesize ** 3))
                   for sc in self.scenes)
        self.pair_cap = int(pair_cap or max(4096, int(cand) + n_wall_tests))
        self.contact_cap = int(contact_cap or max(1024, 8 * n_mov))
        dev = self.device

        def cat(key, dtype):
            parts = []
            for sc in self.scenes:
                a = getattr(sc, key)
                if key == "verts" and a.shape[1] < self.vmax:
                    a = np.concatenate([a, np.zeros((a.shape[0], self.vmax - a.shape[1], 3))], axis=1)
                parts.append(a)
            return torch.from_numpy(np.ascontiguousarray(np.concatenate(parts), dtype=dtype)).to(dev)

        self.t = {}
        self.t["com"] = cat("com", np.float64)
        self.t["verts"] = cat("verts", np.float64)
        for k in ("v", "v12", "w", "wbf", "w12bf", "dxall"):
            self.t[k] = torch.zeros((self.n_rows, 3), dtype=torch.float64, device=dev)
        for k in ("q", "q12", "qall"):
            q = torch.zeros((self.n_rows, 4), dtype=torch.float64, device=dev)
            q[:, 0] = 1.0
            self.t[k] = q
        self.t["nverts"] = cat("nverts", np.int32)
        for k in ("mass", "I", "Iinv", "isize", "fext", "text", "wall_n", "wall_c"):
            self.t[k] = cat(k, np.float64)

        sc_arr = (_lib.Scene * len(self.scenes))()
        for s, sc in enumerate(self.scenes):
            d = sc_arr[s]
            d.row0 = int(self.row0[s])
            d.n_walls = sc.n_walls
            d.n_active = sc.n_rows
            for k in range(6):
                d.box[k] = float(sc.box[k])
            for k in range(3):
                d.cell[k] = float(sc.cell[k])
            d.ncx, d.ncy, d.kmin, d.kspan = sc.ncx, sc.ncy, sc.kmin, sc.kspan
            d.wall_thr = sc.wall_thr
            d.law_p = _lib.law(sc.law_p.Kn, sc.law_p.en, sc.law_p.et, sc.law_p.mu, self.dt, sc.law_p.mode)
            d.law_w = _lib.law(sc.law_w.Kn, sc.law_w.en, sc.law_w.et, sc.law_w.mu, self.dt, sc.law_w.mode)
        self._scene_arr = sc_arr
        lib = _lib.load()
        ws = C.c_void_p()
        with torch.cuda.device(self.device):
            torch.cuda.synchronize(self.device)
            _lib.check(lib.mwd_workspace_create(C.byref(ws), self.device.index, sc_arr, len(self.scenes),
                                                self.n_rows, self.vmax, self.pair_cap, self.contact_cap, self.dt,
                                                self.stream.cuda_stream or None), "mwd_workspace_create")
        self.ws = ws
        self.sphere_test = sphere_test
        # walls with a prescribed velocity (their rows of v) move every step
        if any(sc.moving_walls for sc in self.scenes):
            for s, sc in enumerate(self.scenes):
                if sc.wall_v is not None:
                    r0, W = int(self.row0[s]), sc.n_walls
                    self.t["v"][r0:r0 + W] = torch.from_numpy(np.ascontiguousarray(sc.wall_v[:W])).to(dev)
            _lib.check(lib.mwd_set_wall_motion(ws, 1), "mwd_set_wall_motion")
        if reduce_fp32:
            _lib.check(lib.mwd_set_reduction(ws, _lib.REDUCE_FP32), "mwd_set_reduction")
        if persistent_history:  # keep each contact's tangential displacement (the reference drops it, Q1)
            _lib.check(lib.mwd_set_tangential_history(ws, _lib.HISTORY_PERSISTENT), "mwd_set_tangential_history")
        if spawn is not None:  # rows join on the device's schedule (scene.Spawn; one scene)
            _lib.check(lib.mwd_set_spawn(ws, int(self.row0[0]) + int(spawn.n_active), int(spawn.every),
                                         int(spawn.per)), "mwd_set_spawn")
        self.spawn = spawn
        self.t["radius"] = torch.zeros(self.n_rows, dtype=torch.float64, device=dev)
        self.state = _lib.State()
        self._fill_state()
        self.refresh_radius()
        self.steps_done = 0

    # ------------------------------------------------------------ plumbing
    def _fill_state(self):
        s = self.state
        s.n_rows = self.n_rows
        s.vmax = self.vmax
        for k in STATE_KEYS + CONST_KEYS:
            setattr(s, k, self.t[k].data_ptr())
        s.radius = self.t["radius"].data_ptr() if self.sphere_test else None

    def refresh_radius(self):
        """Recompute the bounding radii from the current vertices (after set_state)."""
        self._enter()
        _lib.check(_lib.load().mwd_radius_init(self.t["com"].data_ptr(), self.t["verts"].data_ptr(),
                                               self.t["nverts"].data_ptr(), self.n_rows, self.vmax,
                                               self.t["radius"].data_ptr(), self.stream.cuda_stream or None),
                   "mwd_radius_init")

    def close(self):
        if getattr(self, "ws", None):
            # work still queued on the stepper's stream must finish before the
            # workspace and (once this object goes) the state tensors are freed
            self.stream.synchronize()
            _lib.load().mwd_workspace_destroy(self.ws)
            self.ws = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --------------------------------------------------------------- stages
    def candidates(self):
        _lib.check(_lib.load().mwd_candidates(self.ws, C.byref(self.state), None), "mwd_candidates")

    def fine_contact(self):
        _lib.check(_lib.load().mwd_fine_contact(self.ws, C.byref(self.state)), "mwd_fine_contact")

    def reduce_verlet(self):
        _lib.check(_lib.load().mwd_reduce_verlet(self.ws, C.byref(self.state)), "mwd_reduce_verlet")

    def _enter(self):
        self.stream.wait_stream(torch.cuda.current_stream(self.device))

    def step(self, n: int = 1):
        """n steps, launched eagerly on the stepper's stream."""
        lib = _lib.load()
        self._enter()
        for _ in range(n):
            _lib.check(lib.mwd_step(self.ws, C.byref(self.state)), "mwd_step")
        self.steps_done += n

    def capture(self, steps_per_graph: int = 10, tangent: bool = False):
        """Capture steps_per_graph consecutive steps (tangent steps if asked)
        in one HIP graph."""
        self._enter()
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        lib = _lib.load()
        with torch.cuda.graph(g, stream=self.stream):
            for _ in range(steps_per_graph):
                if tangent:
                    _lib.check(lib.mwd_step_tangent(self.ws, C.byref(self.state), self._tan_ptrs),
                               "mwd_step_tangent")
                else:
                    _lib.check(lib.mwd_step(self.ws, C.byref(self.state)), "mwd_step")
        torch.cuda.synchronize(self.device)
        self._graph = g
        self._graph_steps = steps_per_graph
        self._graph_tangent = tangent
        return g

    def run(self, n: int):
        """n steps: graph replays for the bulk (when captured), eager steps for
        the rest (tangent steps when the graph holds tangent steps)."""
        g = getattr(self, "_graph", None)
        done = 0
        if g is not None:
            self._enter()
            with torch.cuda.stream(self.stream):
                while n - done >= self._graph_steps:
                    g.replay()
                    done += self._graph_steps
            self.steps_done += done
        if n > done:
            if getattr(self, "_graph_tangent", False):
                self.step_tangent(n - done)
            else:
                self.step(n - done)

    def tuning(self) -> dict:
        """Launch choices of the workspace (cell_pass, gjk_group, gjk_blocks, sphere_block, verlet_block,
        cell_items, verlet_waves, vertex_pack, epa_blocks, epa_fused, fill_cells,
        gjk_buckets, fill_split)."""
        t = _lib.Tuning()
        _lib.check(_lib.load().mwd_get_tuning(self.ws, C.byref(t)), "mwd_get_tuning")
        return {name: int(getattr(t, name)) for name, _ in _lib.Tuning._fields_}

    def set_tuning(self, **choices):
        """Change launch choices for later steps; results are unchanged bit for
        bit. Graphs captured before keep their launches, so they are dropped."""
        cur = self.tuning()
        unknown = set(choices) - set(cur)
        if unknown:
            raise ValueError(f"unknown tuning keys {sorted(unknown)}")
        cur.update(choices)
        _lib.check(_lib.load().mwd_set_tuning(self.ws, C.byref(_lib.Tuning(**cur))), "mwd_set_tuning")
        self._graph = None

    def set_timing(self, on: bool = True):
        _lib.check(_lib.load().mwd_set_timing(self.ws, 1 if on else 0), "mwd_set_timing")

    def kernel_times(self) -> dict:
        """Accumulated kernel-only milliseconds of the single-kernel stages since
        the last call (timing on; call before stage_times)."""
        ms = (C.c_double * _lib.N_KTIMED)()
        n = C.c_int64()
        _lib.check(_lib.load().mwd_kernel_times(self.ws, ms, C.byref(n)), "mwd_kernel_times")
        return dict(steps=n.value, ms=dict(zip(_lib.KTIMED_NAMES, list(ms))))

    def stage_times(self) -> dict:
        """Accumulated per-stage milliseconds since the last call (timing on)."""
        ms = (C.c_double * _lib.N_STAGES)()
        n = C.c_int64()
        _lib.check(_lib.load().mwd_stage_times(self.ws, ms, C.byref(n)), "mwd_stage_times")
        return dict(steps=n.value, ms=dict(zip(_lib.STAGE_NAMES, list(ms))))

    # ------------------------------------------------------------ tangents
    def set_law_seeds(self, seed=None):
        """Give the workspace the law coefficients' derivatives with respect to
        three parameters theta_k: seed[i][k] = d(input_i)/d(theta_k) for the
        law inputs (Kn, en, et, mu), on both laws of every scene; the default
        theta is (Kn, en, mu). The tangent and the adjoint steps read them."""
        lib = _lib.load()
        if seed is None:
            seed = [[1, 0, 0], [0, 1, 0], [0, 0, 0], [0, 0, 1]]
        tans = (_lib.LawTan * (2 * len(self.scenes)))()
        for s, sc in enumerate(self.scenes):
            for j, L in enumerate((sc.law_p, sc.law_w)):
                tans[2 * s + j] = _lib.law_tangent(L.Kn, L.en, L.et, L.mu, seed)
        self._enter()
        torch.cuda.synchronize(self.device)
        _lib.check(lib.mwd_set_law_tangents(self.ws, tans), "mwd_set_law_tangents")

    def step_adjoint(self, lam_ptrs, dtheta: torch.Tensor):
        """Reverse the step whose input state the stepper holds: lam_ptrs (11
        device pointers, state layout) go from the adjoint of the step's output
        to that of its input; dtheta (device, 3) accumulates d/d theta."""
        self._enter()
        _lib.check(_lib.load().mwd_step_adjoint(self.ws, C.byref(self.state), lam_ptrs, dtheta.data_ptr()),
                   "mwd_step_adjoint")

    def capture_adjoint(self, lam_ptrs, dtheta: torch.Tensor, taped: bool = False, steps: int = 1):
        """A HIP graph of `steps` step_adjoint calls (step_adjoint_taped if
        taped) on these adjoint arrays (capture runs nothing). step_adjoint or
        tape_init must have run before (they allocate)."""
        self._enter()
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        lib = _lib.load()
        fn = "mwd_step_adjoint_taped" if taped else "mwd_step_adjoint"
        with torch.cuda.graph(g, stream=self.stream):
            for _ in range(steps):
                _lib.check(getattr(lib, fn)(self.ws, C.byref(self.state), lam_ptrs, dtheta.data_ptr()), fn)
        torch.cuda.synchronize(self.device)
        return g

    # ------------------------------------------------------------ the tape
    def tape_init(self, n_slots: int) -> int:
        """Room for n_slots recorded steps on the reverse sweep's tape (0
        frees it); returns the bytes of one slot. The tape starts empty."""
        self._enter()
        b = C.c_int64()
        _lib.check(_lib.load().mwd_tape_init(self.ws, int(n_slots), C.byref(b)), "mwd_tape_init")
        return int(b.value)

    def tape_rewind(self):
        """Empty the tape (stream-ordered, on the device)."""
        self._enter()
        _lib.check(_lib.load().mwd_tape_rewind(self.ws), "mwd_tape_rewind")

    def step_record(self, n: int = 1):
        """n steps (values as step()), each pushing its contact-stage outputs
        onto the tape."""
        lib = _lib.load()
        self._enter()
        for _ in range(n):
            _lib.check(lib.mwd_step_record(self.ws, C.byref(self.state)), "mwd_step_record")
        self.steps_done += n

    def capture_record(self, steps: int = 1):
        """A HIP graph of `steps` step_record calls (capture runs nothing)."""
        self._enter()
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        lib = _lib.load()
        with torch.cuda.graph(g, stream=self.stream):
            for _ in range(steps):
                _lib.check(lib.mwd_step_record(self.ws, C.byref(self.state)), "mwd_step_record")
        torch.cuda.synchronize(self.device)
        return g

    def step_adjoint_taped(self, lam_ptrs, dtheta: torch.Tensor):
        """step_adjoint of the step last recorded on the tape: its input state
        and contact-stage outputs are taken back from the tape (the stepper's
        state becomes that input state)."""
        self._enter()
        _lib.check(_lib.load().mwd_step_adjoint_taped(self.ws, C.byref(self.state), lam_ptrs, dtheta.data_ptr()),
                   "mwd_step_adjoint_taped")

    def enable_tangents(self, seed=None):
        """Carry d(state)/d(theta) for three law parameters theta_k.

        seed[i][k] = d(input_i)/d(theta_k) for the law inputs (Kn, en, et, mu),
        applied to both laws (particle and wall) of every scene. The default
        theta is (Kn, en, mu), with et held fixed. Tangents start at zero: the
        initial state does not depend on theta."""
        self.set_law_seeds(seed)
        self.tan = {k: torch.zeros((_lib.N_TAN,) + tuple(self.t[k].shape), dtype=torch.float64, device=self.device)
                    for k in _lib.TAN_ARRAYS}
        arr = (C.c_void_p * (len(_lib.TAN_ARRAYS) * _lib.N_TAN))()
        for a, k in enumerate(_lib.TAN_ARRAYS):
            base, n = self.tan[k].data_ptr(), self.tan[k][0].numel()
            for j in range(_lib.N_TAN):
                arr[_lib.N_TAN * a + j] = base + 8 * n * j
        self._tan_ptrs = arr

    def numpy_tangents(self) -> dict:
        """The tangent arrays [N_TAN, ...] as numpy (after the stepper's stream
        has finished)."""
        torch.cuda.synchronize(self.device)
        return {k: v.cpu().numpy().copy() for k, v in self.tan.items()}

    def reset_tangents(self):
        for v in self.tan.values():
            v.zero_()

    def step_tangent(self, n: int = 1):
        """n steps that also advance the tangents (values identical to step())."""
        lib = _lib.load()
        self._enter()
        for _ in range(n):
            _lib.check(lib.mwd_step_tangent(self.ws, C.byref(self.state), self._tan_ptrs), "mwd_step_tangent")
        self.steps_done += n

    # --------------------------------------------------------------- status
    def status(self) -> dict:
        st = _lib.Status()
        _lib.check(_lib.load().mwd_step_status(self.ws, C.byref(st)), "mwd_step_status")
        return dict(n_pairs=st.n_pairs, n_wall_pairs=st.n_wall_pairs, n_contacts=st.n_contacts,
                    n_cells=st.n_cells, flags=st.flags, n_sphere_pass=st.n_sphere_pass)

    def check(self) -> dict:
        s = self.status()
        f = s["flags"]
        if f & _lib.F_PAIR_OVERFLOW:
            raise _lib.MwdError(_lib.MWD_E_CAPACITY, "step", f"{s['n_pairs']} candidate pairs > pair_cap {self.pair_cap}")
        if f & _lib.F_CONTACT_OVERFLOW:
            raise _lib.MwdError(_lib.MWD_E_CAPACITY, "step", f"{s['n_contacts']} contacts > contact_cap {self.contact_cap}")
        if f & _lib.F_KEYRANGE:
            raise _lib.MwdError(_lib.MWD_E_KEYRANGE, "step", "a cell key left its scene's key range")
        if f & _lib.F_CELL:
            raise _lib.MwdError(_lib.MWD_E_CELL, "step", f"a cell holds more than {_lib.MAX_CELL} particles")
        if f & _lib.F_ROW_CONTACTS:
            raise _lib.MwdError(_lib.MWD_E_CAPACITY, "step",
                                "the contact-record overflow list is full; rerun with a larger contact_cap")
        if f & _lib.F_TAPE:
            raise _lib.MwdError(_lib.MWD_E_CAPACITY, "step",
                                "a recorded step found the tape full, or a taped adjoint step found it empty")
        return s

    # ---------------------------------------------------------------- views
    def pairs(self) -> np.ndarray:
        p = C.c_void_p()
        npairs = self.status()["n_pairs"]
        _lib.check(_lib.load().mwd_pairs_view(self.ws, C.byref(p), None), "mwd_pairs_view")
        n = min(npairs, self.pair_cap)
        return _copy_from_device(p.value, n * 2, np.int32).reshape(n, 2)

    def contacts(self) -> dict:
        """This step's contacts in candidate-pair order: pair index, EPA depth,
        normal and contact points, the law's force on A and torques, EPA
        status; plus the per-pair collision codes."""
        ptrs = [C.c_void_p() for _ in range(11)]
        _lib.check(_lib.load().mwd_contacts_view(self.ws, *[C.byref(x) for x in ptrs]), "mwd_contacts_view")
        s = self.status()
        P = min(s["n_pairs"], self.pair_cap)
        K = min(s["n_contacts"], self.contact_cap)
        coll = _copy_from_device(ptrs[0].value, P, np.int32)
        out = dict(collision=coll, pair=_copy_from_device(ptrs[1].value, K, np.int32))
        for i, name in ((2, "depth"),):
            out[name] = _copy_from_device(ptrs[i].value, K, np.float64)
        for i, name in ((3, "normal"), (4, "cpA"), (5, "cpB"), (6, "Fn"), (7, "tauA"), (8, "tauB")):
            out[name] = _copy_from_device(ptrs[i].value, 3 * K, np.float64).reshape(K, 3)
        out["epa_status"] = _copy_from_device(ptrs[9].value, K, np.int32)
        # contacts are stored in the order GJK found them; present them in
        # candidate-pair order
        order = np.argsort(out["pair"], kind="stable")
        for k in ("pair", "depth", "normal", "cpA", "cpB", "Fn", "tauA", "tauB", "epa_status"):
            out[k] = out[k][order]
        return out

    def numpy_state(self) -> dict:
        torch.cuda.synchronize(self.device)
        return {k: self.t[k].cpu().numpy().copy() for k in STATE_KEYS}

    def scene_rows(self, s: int) -> slice:
        return slice(int(self.row0[s]), int(self.row0[s]) + self.scenes[s].n_rows)

    def snapshot(self) -> dict:
        """Device copies of the state tensors (and radii), for restore()."""
        self._enter()
        with torch.cuda.stream(self.stream):
            snap = {k: self.t[k].clone() for k in STATE_KEYS + ("radius", "wall_c")}
        return snap

    def restore(self, snap: dict):
        """Put back the state of a snapshot() (stream-ordered, on the device)."""
        self._enter()
        with torch.cuda.stream(self.stream):
            for k, v in snap.items():
                self.t[k].copy_(v)

    def set_state(self, st: dict):
        """Overwrite state tensors from numpy arrays (rows of all scenes)."""
        torch.cuda.synchronize(self.device)
        for k, a in st.items():
            self.t[k].copy_(torch.from_numpy(np.ascontiguousarray(a)).to(self.device))
        torch.cuda.synchronize(self.device)
        self.refresh_radius()


# ------------------------------------------------------------ raw copies
_HIP = None


def _hip():
    """The HIP runtime already loaded by torch (matched by its soname, so no
    second runtime is mapped into the process)."""
    global _HIP
    if _HIP is None:
        import ctypes
        _HIP = ctypes.CDLL("libamdhip64.so.7")
        _HIP.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _HIP.hipMemcpy.restype = ctypes.c_int
    return _HIP


def _check_hip(rc):
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed ({rc})")


def _copy_from_device(addr: int, n: int, dtype) -> np.ndarray:
    out = np.empty(n, dtype=dtype)
    if n:
        torch.cuda.synchronize()
        import ctypes
        _check_hip(_hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(addr),
                                    ctypes.c_size_t(out.nbytes), 2))  # hipMemcpyDeviceToHost
    return out

"""On-disk outputs of the reference driver, written from device state.

* Legacy-VTK POLYDATA snapshots in the layout of LIBRARY_Vtk_creator.py:3-41:
  a four-line header, "POINTS\\t<n>\\tfloat", one "x\\ty\\tz" line per hull
  vertex of rows [row_min, row_max), "POLYGONS\\t<f>\\t<4f>" and one
  "3\\ta\\tb\\tc" line per triangle, vertex ids offset by the vertices of the
  rows before. Coordinates are the shortest round-trip decimal form of the
  doubles (Python's str of a float64, as the reference writes them).
* Per-particle time series Data_<id>.dat in the layout of
  LIBRARY_Write_particle_data.py:3-12: a header line at timestep 0 (file
  truncated), then "timestep, com, vx, vy, |vz|, force_normal, force_n_spring,
  force_n_damping" tab-separated. The reference's driver never sets the three
  force vectors (MWDiem.py:732-734 and :1015 are commented out), so they are
  written as its initial integer zeros ("0").

Triangles are the scene's (SceneData.faces: each hull's triangles as Qhull
listed them when the generator built it, Particle_class.py:72-87, which the
reference writes unchanged); for a scene without them, the convex hull of
each row's vertices, oriented outward, computed once per writer (a rigid
motion leaves the triangulation unchanged).

AsyncWriter copies the state a snapshot needs into pinned host buffers on a
side stream (after the work already queued on the stepper's stream) and
writes the files on a worker thread once the copies have landed, so the GPU
keeps stepping while files are written.
"""
from __future__ import annotations

import queue
import threading
from pathlib import Path

import numpy as np
from scipy.spatial import ConvexHull


def hull_faces(verts: np.ndarray, nverts: np.ndarray) -> list:
    """Outward-oriented triangles (vertex indices into the row's own vertex
    list) of every row's convex hull."""
    out = []
    for r in range(verts.shape[0]):
        V = verts[r, :int(nverts[r])]
        tri = ConvexHull(V).simplices.astype(np.int64)
        cm = V.mean(axis=0)
        d0, d1, d2 = V[tri[:, 0]], V[tri[:, 1]], V[tri[:, 2]]
        flip = np.einsum("ij,ij->i", np.cross(d1 - d0, d2 - d1), d0 - cm) < 0
        tri[flip] = tri[flip][:, [0, 2, 1]]
        out.append(tri)
    return out


def vtk_text(verts: np.ndarray, nverts: np.ndarray, faces: list, row_min: int, row_max: int) -> str:
    rows = range(row_min, row_max)
    lines = ["# vtk DataFile Version 2.0", "# kata", "ASCII", "DATASET POLYDATA"]
    n_pts = int(sum(int(nverts[r]) for r in rows))
    lines.append(f"POINTS\t{n_pts}\tfloat")
    for r in rows:
        for x, y, z in verts[r, :int(nverts[r])].tolist():
            lines.append(f"{x!r}\t{y!r}\t{z!r}")
    n_f = int(sum(len(faces[r]) for r in rows))
    lines.append(f"POLYGONS\t{n_f}\t{n_f * 4}")
    addon = 0
    for r in rows:
        for a, b, c in (faces[r] + addon).tolist():
            lines.append(f"3\t{a}\t{b}\t{c}")
        addon += int(nverts[r])
    return "\n".join(lines) + "\n"


def write_vtk(path, verts, nverts, faces, row_min: int, row_max: int) -> None:
    Path(path).write_text(vtk_text(verts, nverts, faces, row_min, row_max))


DATA_HEADER = "Timestep\t Cm_x\t Cm_y\t Cm_z\t vx\t vy\t vz\t fx\t fy\t fz\t fsx\t fsy\t fsz\t fdx\t fdy\t fdz\n"


def data_line(timestep: int, com, v) -> str:
    c = [float(x) for x in com]
    u = [float(x) for x in v]
    return (f"{timestep}\t{c[0]!r}\t{c[1]!r}\t{c[2]!r}\t{u[0]!r}\t{u[1]!r}\t{abs(u[2])!r}"
            + "\t0" * 9 + "\n")


def write_particle_data(prefix: str, row_ids, timestep: int, com: np.ndarray, v: np.ndarray) -> None:
    """Append (truncate at timestep 0) one line per row to <prefix><id>.dat."""
    for r in row_ids:
        mode = "w" if timestep == 0 else "a"
        with open(f"{prefix}{int(r)}.dat", mode) as f:
            if timestep == 0:
                f.write(DATA_HEADER)
            f.write(data_line(timestep, com[r], v[r]))


class AsyncWriter:
    """Outputs of a Stepper's first scene (rows of all scenes when it holds
    several), written off the critical path.

    vtk(step) writes <dirname>/Out_<step>.vtk of rows [0, n_rows) (the
    reference writes all rows, walls included, MWDiem.py:1060); data(step)
    appends the moving rows' lines to <data_prefix><id>.dat. close() waits
    for the files."""

    def __init__(self, S, dirname, data_prefix: str | None = None):
        import torch
        self.S = S
        self.dir = Path(dirname)
        self.dir.mkdir(parents=True, exist_ok=True)
        self.prefix = data_prefix if data_prefix is not None else str(self.dir / "Data_")
        self.side = torch.cuda.Stream(device=S.device)  # default priority
        nv = S.t["nverts"].cpu().numpy()
        self.nverts = nv
        if all(sc.faces is not None for sc in S.scenes):
            self.faces = [np.asarray(f, dtype=np.int64) for sc in S.scenes for f in sc.faces]
        else:
            self.faces = hull_faces(S.t["verts"].cpu().numpy(), nv)
        self.W = S.scenes[0].n_walls
        self.jobs: queue.Queue = queue.Queue()
        self.error = None
        self.worker = threading.Thread(target=self._run, daemon=True)
        self.worker.start()

    def _snapshot(self, keys):
        """Device copies of the state arrays taken in the stepper's stream
        order (so the next step cannot overwrite them first), then copied to
        pinned host buffers on the side stream while the stepper goes on."""
        import torch
        ev = torch.cuda.Event()
        with torch.cuda.stream(self.S.stream):
            staged = {k: self.S.t[k].clone() for k in keys}
        self.side.wait_stream(self.S.stream)
        host = {}
        with torch.cuda.stream(self.side):
            for k, t in staged.items():
                t.record_stream(self.side)  # its memory is reused only after the side stream's copy
                h = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
                h.copy_(t, non_blocking=True)
                host[k] = h
            ev.record(self.side)
        return host, ev

    def vtk(self, step: int) -> None:
        host, ev = self._snapshot(("verts",))
        self.jobs.put(("vtk", step, host, ev))

    def data(self, step: int) -> None:
        host, ev = self._snapshot(("com", "v"))
        self.jobs.put(("data", step, host, ev))

    def _run(self):
        while True:
            job = self.jobs.get()
            if job is None:
                return
            kind, step, host, ev = job
            try:
                ev.synchronize()
                if kind == "vtk":
                    V = host["verts"].numpy()
                    write_vtk(self.dir / f"Out_{step}.vtk", V, self.nverts, self.faces, 0, V.shape[0])
                else:
                    com, v = host["com"].numpy(), host["v"].numpy()
                    write_particle_data(self.prefix, range(self.W, com.shape[0]), step, com, v)
            except Exception as e:  # reported by close()
                self.error = e

    def close(self) -> None:
        self.jobs.put(None)
        self.worker.join()
        if self.error is not None:
            raise self.error

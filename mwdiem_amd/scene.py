"""Scene generator: convex polyhedra in a walled box, in the reference's data layout.

A scene is what the reference driver packs into its SoA arrays before the time
loop (MWDiem.py:603-647): rows [0, n_walls) are wall boxes, the remaining rows
are particles; world vertices [N, vmax, 3] zero-padded; centre of mass; mass;
body-frame inertia (diagonal) and its inverse; the inertia "size"
(trace/3, Particle_class.py:151); external force m*g.

Particle geometry follows Particle_class.Spherical_convex_polygon_maker_3D
(CLASSES/Particle_class.py:239-303): the convex hull of random surface points;
the centre of mass is the mean of the hull vertices (not the volume centroid);
the inertia tensor about it is computed with Dobrovolskis' polyhedron
formulae (:125-156); I_bf = diag(eigenvalues) in LAPACK order while the
vertices stay in the hull frame (:288-291). Points are drawn area-uniformly on
an ellipsoid, each particle gets a uniformly random orientation, and particles
are seeded without overlap on a jittered body-centred-cubic lattice.

Units: mm, kg, s.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
from scipy.spatial import ConvexHull

G_MM = 9810.0          # mm/s^2
RHO_KG_MM3 = 2.5e-6    # 2500 kg/m^3


REF = 0       # the reference's contact law, quirks included
PHYSICAL = 1  # midpoint contact point, effective-mass damping, tangential force on linear momentum


@dataclass
class Law4:
    Kn: float = 1e4
    en: float = 0.5
    et: float = 0.5
    mu: float = 0.3
    mode: int = REF


@dataclass
class SceneData:
    n_walls: int
    vmax: int
    com: np.ndarray
    verts: np.ndarray
    nverts: np.ndarray
    mass: np.ndarray
    I: np.ndarray
    Iinv: np.ndarray
    isize: np.ndarray
    fext: np.ndarray
    text: np.ndarray
    wall_n: np.ndarray
    wall_c: np.ndarray
    box: np.ndarray            # x_max, x_min, y_max, y_min, z_max, z_min (MWDiem.py:553-572)
    cell: np.ndarray           # cellsize x, y, z (MWDiem.py:578-580)
    ncx: int
    ncy: int
    ncz: int
    kmin: int
    kspan: int
    max_particlesize: float
    law_p: Law4 = field(default_factory=Law4)
    law_w: Law4 = field(default_factory=Law4)
    dt: float = 1e-6
    name: str = ""
    inner: tuple = (0.0, 0.0, 0.0)   # inner box extent (Lx, Ly, Lz)
    wall_v: np.ndarray | None = None  # [N,3] prescribed wall velocities (walls' rows); None: static walls
    pairs_per_particle: float | None = None  # candidate pairs to provide for per particle (None: the stepper's default)
    faces: list | None = None  # per row, the hull's triangles in Qhull order (None: the writer takes the hull's)

    @property
    def moving_walls(self) -> bool:
        return self.wall_v is not None and bool(np.any(self.wall_v[:self.n_walls] != 0.0))

    @property
    def n_rows(self) -> int:
        return self.com.shape[0]

    @property
    def n_particles(self) -> int:
        return self.n_rows - self.n_walls

    @property
    def wall_thr(self) -> float:
        return 6 * self.max_particlesize

    def copy(self) -> "SceneData":
        import copy
        return copy.deepcopy(self)

    _ARRAYS = ("com", "verts", "nverts", "mass", "I", "Iinv", "isize", "fext", "text", "wall_n", "wall_c", "box",
               "cell")

    def to_arrays(self) -> dict:
        """Plain arrays (npz-friendly) that from_arrays() turns back into the scene."""
        d = {k: getattr(self, k) for k in self._ARRAYS}
        d["ints"] = np.array([self.n_walls, self.vmax, self.ncx, self.ncy, self.ncz, self.kmin, self.kspan], np.int64)
        d["floats"] = np.array([self.max_particlesize, self.dt] + list(self.inner))
        d["laws"] = np.array([[L.Kn, L.en, L.et, L.mu, L.mode] for L in (self.law_p, self.law_w)], float)
        return d

    @classmethod
    def from_arrays(cls, d) -> "SceneData":
        i = [int(x) for x in d["ints"]]
        f = [float(x) for x in d["floats"]]
        laws = [Law4(Kn=float(r[0]), en=float(r[1]), et=float(r[2]), mu=float(r[3]), mode=int(r[4]))
                for r in np.asarray(d["laws"])]
        kw = {k: np.array(d[k]) for k in cls._ARRAYS}
        return cls(n_walls=i[0], vmax=i[1], ncx=i[2], ncy=i[3], ncz=i[4], kmin=i[5], kspan=i[6],
                   max_particlesize=f[0], dt=f[1], inner=tuple(f[2:5]), law_p=laws[0], law_w=laws[1], **kw)


# ----------------------------------------------------------------- geometry
def ellipsoid_points(rng: np.random.Generator, n: int, axes) -> np.ndarray:
    """n points distributed uniformly by area on the ellipsoid with semi-axes `axes`."""
    a, b, c = axes
    gmax = max(b * c, a * c, a * b)
    out = np.empty((0, 3))
    while out.shape[0] < n:
        m = 4 * (n - out.shape[0]) + 8
        u = rng.normal(size=(m, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        g = np.sqrt((b * c * u[:, 0]) ** 2 + (a * c * u[:, 1]) ** 2 + (a * b * u[:, 2]) ** 2)
        keep = rng.random(m) * gmax < g
        out = np.vstack([out, u[keep] * np.array([a, b, c])])
    return out[:n]


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ])


def hull_body(points: np.ndarray, qhull_faces: bool = False):
    """Hull vertices (in scipy's vertex order), outward-oriented triangles as
    vertex indices, and the hull volume. With qhull_faces, also the triangles
    as Qhull lists them (facet order, each facet's own vertex order), mapped
    to the vertex indices: the face ids of Particle_class.py:72-87, which the
    VTK writer prints."""
    h = ConvexHull(points)
    vidx = h.vertices
    remap = -np.ones(points.shape[0], dtype=np.int64)
    remap[vidx] = np.arange(vidx.size)
    tri = remap[h.simplices]
    qtri = tri.astype(np.int32)
    V = points[vidx]
    # orient each face outward w.r.t. the vertex mean (Particle_class.py:109-113)
    cm = V.mean(axis=0)
    d0, d1, d2 = V[tri[:, 0]], V[tri[:, 1]], V[tri[:, 2]]
    nrm = np.cross(d1 - d0, d2 - d1)
    flip = np.einsum("ij,ij->i", nrm, d0 - cm) < 0
    tri[flip] = tri[flip][:, [0, 2, 1]]
    if qhull_faces:
        return V, tri, float(h.volume), qtri
    return V, tri, float(h.volume)


def polyhedron_inertia(V: np.ndarray, tri: np.ndarray, density: float, reference_volume_element: bool = False):
    """Inertia tensor about the origin of a closed triangulated polyhedron
    (Dobrovolskis 1996, the formulae of Particle_class.Calc_inertia :125-156).

    With reference_volume_element=True the per-face volume element is the
    reference's D.n_unit/6 (n normalised, Particle_class.py:110,137) instead of
    the tetrahedron volume D.(E x F)/6.
    """
    D, E, F = V[tri[:, 0]], V[tri[:, 1]], V[tri[:, 2]]
    if reference_volume_element:
        n = np.cross(E - D, F - E)
        n /= np.linalg.norm(n, axis=1, keepdims=True)
        dV = np.einsum("ij,ij->i", D, n) / 6
    else:
        dV = np.einsum("ij,ij->i", D, np.cross(E, F)) / 6

    def P(i, j):
        return np.sum(density * dV / 20 * (2 * D[:, i] * D[:, j] + 2 * E[:, i] * E[:, j] + 2 * F[:, i] * F[:, j]
                                           + D[:, i] * E[:, j] + D[:, j] * E[:, i] + D[:, i] * F[:, j]
                                           + D[:, j] * F[:, i] + E[:, i] * F[:, j] + E[:, j] * F[:, i]))
    Pxx, Pyy, Pzz = P(0, 0), P(1, 1), P(2, 2)
    Pxy, Pxz, Pyz = P(0, 1), P(0, 2), P(1, 2)
    I = np.array([[Pyy + Pzz, -Pxy, -Pxz], [-Pxy, Pxx + Pzz, -Pyz], [-Pxz, -Pyz, Pxx + Pyy]])
    return I, (I[0, 0] + I[1, 1] + I[2, 2]) / 3


def body_frame(I: np.ndarray):
    """I_bf = diag(eig(I)) in LAPACK order and its inverse (Particle_class.py:289-291)."""
    ev = np.linalg.eig(I)[0].real
    Ibf = np.diag(ev)
    return Ibf, np.linalg.inv(Ibf)


def make_particle(rng, npts: int, axes, size: float, density: float, ref_inertia: bool, consistent: bool = False):
    """One hull. With consistent=False the inertia follows the reference
    (diag(eig(I)) while the vertices stay in the hull frame, and the reference's
    volume element when ref_inertia); with consistent=True the full tensor is
    expressed in the world frame of the rotated vertices (the frame in which
    q = 1 at t = 0), so geometry and inertia agree."""
    pts = ellipsoid_points(rng, npts, np.asarray(axes) * size)
    V, tri, vol, qtri = hull_body(pts, qhull_faces=True)
    cm = V.mean(axis=0)
    V = V - cm
    I, isize = polyhedron_inertia(V, tri, density, ref_inertia)
    R = random_rotation(rng)
    Vw = V @ R.T
    if consistent:
        Iw = R @ I @ R.T
        Ibf, Iinv = Iw, np.linalg.inv(Iw)
    else:
        Ibf, Iinv = body_frame(I)
    return Vw, density * vol, Ibf, Iinv, isize, face_normals(Vw, tri), qtri


_CHUNK = 2048


def draw_particles(n: int, seed: int, npts, axes, size):
    """The generator's random draws alone, in make_particles' order (chunk c
    from PCG64([seed, c]); per particle its point count, size factor, points
    and rotation): a list of (points, rotation). make_particles and
    make_particles_device build hulls from exactly these draws."""
    out = []
    for c in range((n + _CHUNK - 1) // _CHUNK):
        rng = np.random.Generator(np.random.PCG64([seed, c]))
        for _ in range(min(_CHUNK, n - c * _CHUNK)):
            k = int(rng.integers(npts[0], npts[1] + 1))
            s = float(rng.uniform(size[0], size[1]))
            pts = ellipsoid_points(rng, k, np.asarray(axes) * s)
            out.append((pts, random_rotation(rng)))
    return out


def hull_bodies_device(points: list, density: float, device: int = 0):
    """Hull bodies of point sets on the GPU (mwd_hull_bodies): per set its
    hull vertices in input order centred on their mean, the volume and the
    inertia tensor about the mean (volume element D.(E x F)/6), and its size.
    Raises for a set the kernel does not triangulate (four points on a
    supporting plane, or a point count outside 4..64)."""
    import ctypes as C

    import torch

    from . import _lib
    H = len(points)
    pmax = max(4, max(int(p.shape[0]) for p in points))
    P = np.zeros((H, pmax, 3))
    n = np.zeros(H, np.int32)
    for i, p in enumerate(points):
        P[i, :p.shape[0]] = p
        n[i] = p.shape[0]
    dev = torch.device("cuda", device)
    tp = torch.from_numpy(P).to(dev)
    tn = torch.from_numpy(n).to(dev)
    out = {"verts": torch.zeros((H, pmax, 3), dtype=torch.float64, device=dev),
           "nverts": torch.zeros(H, dtype=torch.int32, device=dev),
           "com": torch.zeros((H, 3), dtype=torch.float64, device=dev),
           "vol": torch.zeros(H, dtype=torch.float64, device=dev),
           "I": torch.zeros((H, 3, 3), dtype=torch.float64, device=dev),
           "isize": torch.zeros(H, dtype=torch.float64, device=dev),
           "status": torch.zeros(H, dtype=torch.int32, device=dev)}
    with torch.cuda.device(dev):
        stream = torch.cuda.current_stream(dev).cuda_stream
        _lib.check(_lib.load().mwd_hull_bodies(tp.data_ptr(), tn.data_ptr(), H, pmax, C.c_double(density),
                                               out["verts"].data_ptr(), out["nverts"].data_ptr(),
                                               out["com"].data_ptr(), out["vol"].data_ptr(), out["I"].data_ptr(),
                                               out["isize"].data_ptr(), out["status"].data_ptr(), stream or None),
                   "mwd_hull_bodies")
        torch.cuda.synchronize(dev)
    res = {k: v.cpu().numpy() for k, v in out.items()}
    bad = np.flatnonzero(res["status"] != 0)
    if bad.size:
        raise ValueError(f"mwd_hull_bodies: {bad.size} point sets not triangulated (first {int(bad[0])}, "
                         f"status {int(res['status'][bad[0]])})")
    return res


def make_particles_device(n: int, seed: int, npts, axes, size, density, consistent=False, device: int = 0):
    """make_particles with the hulls, volumes and inertia tensors built on the
    GPU in one launch (the draws are make_particles' own). The vertices come
    out in the same order; centres, volumes and tensors agree with the host's
    to rounding (the host's centre is numpy's mean, the volume Qhull's). No
    face lists (the output writer then triangulates each hull), and the
    reference's volume element is not offered."""
    draws = draw_particles(n, seed, npts, axes, size)
    hb = hull_bodies_device([d[0] for d in draws], density, device)
    out = []
    for i, (_, R) in enumerate(draws):
        V = hb["verts"][i, :hb["nverts"][i]]
        I = hb["I"][i]
        Vw = V @ R.T
        if consistent:
            Iw = R @ I @ R.T
            Ibf, Iinv = Iw, np.linalg.inv(Iw)
        else:
            Ibf, Iinv = body_frame(I)
        out.append((Vw, density * float(hb["vol"][i]), Ibf, Iinv, float(hb["isize"][i]), None, None))
    return out


def _particle_chunk(args):
    seed, chunk, n, npts, axes, size, density, ref_inertia, consistent = args
    rng = np.random.Generator(np.random.PCG64([seed, chunk]))
    out = []
    for _ in range(n):
        k = int(rng.integers(npts[0], npts[1] + 1))
        s = float(rng.uniform(size[0], size[1]))
        out.append(make_particle(rng, k, axes, s, density, ref_inertia, consistent))
    return out


def gpu_runtime_started() -> bool:
    """True once this process has initialised the GPU runtime (torch's HIP
    context, or the HIP library loaded by mwdiem_amd._lib). A process in that
    state must not fork workers: the child inherits the runtime's locks and
    device mappings without its threads."""
    import sys
    torch = sys.modules.get("torch")
    if torch is not None and torch.cuda.is_initialized():
        return True
    lib = sys.modules.get("mwdiem_amd._lib")
    return lib is not None and getattr(lib, "_lib", None) is not None


def make_particles(n: int, seed: int, npts, axes, size, density, ref_inertia, consistent=False,
                   workers: int | None = None):
    """n particles in chunks of 2048, chunk c drawn from PCG64([seed, c]): the
    result does not depend on the number of worker processes. Workers are
    forked only while the GPU runtime is not started in this process (after
    that the chunks are made in this process, with the same result)."""
    jobs = [(seed, c, min(_CHUNK, n - c * _CHUNK), tuple(npts), tuple(axes), tuple(size), density, ref_inertia,
             consistent) for c in range((n + _CHUNK - 1) // _CHUNK)]
    if len(jobs) <= 1 or gpu_runtime_started():
        return [p for j in jobs for p in _particle_chunk(j)]
    import multiprocessing as mp
    import os
    nw = workers or min(len(jobs), int(os.environ.get("MWD_GEN_WORKERS", "0")) or min(16, os.cpu_count() or 1))
    if nw <= 1:
        return [p for j in jobs for p in _particle_chunk(j)]
    with mp.get_context("fork").Pool(nw) as pool:
        res = pool.map(_particle_chunk, jobs)
    return [p for r in res for p in r]


def box_corners(lo, hi):
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    return np.array([[x, y, z] for x in (hi[0], lo[0]) for y in (hi[1], lo[1]) for z in (hi[2], lo[2])])


def walls_for_box(L, T: float, density: float, ref_inertia: bool):
    """Six wall boxes of thickness T around [0,Lx]x[0,Ly]x[0,Lz]:
    floor, ceiling, x-, x+, y-, y+. Returns per-wall (vertices, normal,
    centre, mass, I, Iinv, isize, Qhull-order faces)."""
    Lx, Ly, Lz = L
    specs = [
        ((-T, -T, -T), (Lx + T, Ly + T, 0.0), (0.0, 0.0, 1.0)),
        ((-T, -T, Lz), (Lx + T, Ly + T, Lz + T), (0.0, 0.0, 1.0)),
        ((-T, -T, 0.0), (0.0, Ly + T, Lz), (1.0, 0.0, 0.0)),
        ((Lx, -T, 0.0), (Lx + T, Ly + T, Lz), (1.0, 0.0, 0.0)),
        ((0.0, -T, 0.0), (Lx, 0.0, Lz), (0.0, 1.0, 0.0)),
        ((0.0, Ly, 0.0), (Lx, Ly + T, Lz), (0.0, 1.0, 0.0)),
    ]
    out = []
    for lo, hi, n in specs:
        P = box_corners(lo, hi)
        V, tri, vol, qtri = hull_body(P, qhull_faces=True)
        c = V.mean(axis=0)
        I, isize = polyhedron_inertia(V - c, tri, density, ref_inertia)
        Ibf, Iinv = body_frame(I)
        out.append((V, np.array(n), c, density * vol, Ibf, Iinv, isize, qtri))
    return out


def bcc_sites(L, spacing: float, r: float, n: int) -> np.ndarray:
    """First n sites of a body-centred-cubic packing filling [r, L-r] from the
    floor up (nearest neighbours `spacing` apart)."""
    Lx, Ly, Lz = L
    a = spacing
    dz = a / math.sqrt(2.0)
    sites = []
    k = 0
    while len(sites) < n:
        z = r + k * dz
        if z > Lz - r + 1e-12:
            raise ValueError(f"box too small for {n} particles at spacing {a}")
        off = 0.0 if k % 2 == 0 else a / 2
        xs = np.arange(r + off, Lx - r + 1e-12, a)
        ys = np.arange(r + off, Ly - r + 1e-12, a)
        X, Y = np.meshgrid(xs, ys, indexing="ij")
        layer = np.stack([X.ravel(), Y.ravel(), np.full(X.size, z)], axis=1)
        sites.extend(layer.tolist())
        k += 1
    return np.asarray(sites[:n])


def face_normals(V: np.ndarray, tri: np.ndarray) -> np.ndarray:
    n = np.cross(V[tri[:, 1]] - V[tri[:, 0]], V[tri[:, 2]] - V[tri[:, 1]])
    return n / np.linalg.norm(n, axis=1, keepdims=True)


def _separated(VA, NA, VB, NB) -> bool:
    """True when a face normal of A or B separates the two hulls (a
    sufficient test: hulls it cannot separate are treated as overlapping)."""
    N = np.vstack([NA, NB])
    pa = VA @ N.T
    pb = VB @ N.T
    return bool(np.any(pa.max(axis=0) < pb.min(axis=0)) or np.any(pb.max(axis=0) < pa.min(axis=0)))


def deposit_sites(parts_V, parts_N, L, rng, gap: float = 0.02, window: int = 4000) -> np.ndarray:
    """Bottom-up deposition: particle i goes to the lowest free candidate site
    (fine grid, random order inside each layer) where it lies inside the box
    and is separated from every particle already placed."""
    Lx, Ly, Lz = L
    n = len(parts_V)
    rcirc = np.array([np.max(np.linalg.norm(V, axis=1)) for V in parts_V])
    rin = np.array([np.min(np.einsum("ij,ij->i", N, V[np.argmax(V @ N.T, axis=0)]))
                    for V, N in zip(parts_V, parts_N)])
    rin_min = float(np.min(rin))
    h = 0.4 * float(np.mean(rcirc))
    xs = np.arange(h / 2, Lx, h)
    ys = np.arange(h / 2, Ly, h)
    zs = np.arange(h / 2, Lz, h)
    layers = []
    for z in zs:
        X, Y = np.meshgrid(xs, ys, indexing="ij")
        lay = np.stack([X.ravel(), Y.ravel(), np.full(X.size, z)], axis=1)
        layers.append(lay[rng.permutation(lay.shape[0])])
    sites = np.concatenate(layers)
    free = np.ones(sites.shape[0], dtype=bool)
    centers = np.zeros((n, 3))
    start = 0
    for i in range(n):
        V, N = parts_V[i], parts_N[i]
        lo, hi = V.min(axis=0), V.max(axis=0)
        placed = False
        w0 = start
        while not placed and w0 < sites.shape[0]:
            idx = np.arange(w0, min(w0 + window, sites.shape[0]))
            idx = idx[free[idx]]
            c = sites[idx]
            ok = np.all(c + lo >= gap, axis=1) & (c[:, 0] + hi[0] <= Lx - gap) & (c[:, 1] + hi[1] <= Ly - gap) \
                & (c[:, 2] + hi[2] <= Lz - gap)
            if i:
                d = np.linalg.norm(c[:, None, :] - centers[None, :i, :], axis=2)
                ok &= np.all(d >= rin[None, :i] + rin[i], axis=1)
            cand = idx[ok]
            if cand.size:
                C = sites[cand]
                good = np.ones(cand.size, dtype=bool)
                if i:
                    dmin = np.linalg.norm(C[:, None, :] - centers[None, :i, :], axis=2)
                    near = np.nonzero(np.any(dmin < rcirc[:i] + rcirc[i] + gap, axis=0))[0]
                    pA = V @ N.T
                    for j in near:
                        Nij = np.vstack([N, parts_N[j]])
                        pa = V @ Nij.T                       # (nv, M)
                        pb = (parts_V[j] + centers[j]) @ Nij.T
                        shift = C @ Nij.T                   # (K, M)
                        sep = np.any((pa.max(axis=0) + shift < pb.min(axis=0)) |
                                     (pb.max(axis=0) < pa.min(axis=0) + shift), axis=1)
                        far = dmin[:, j] >= rcirc[j] + rcirc[i] + gap
                        good &= sep | far
                    del pA
                hit = np.nonzero(good)[0]
                if hit.size:
                    k = cand[hit[0]]
                    ck = sites[k]
                    centers[i] = ck
                    # sites inside this particle's insphere can host no centre
                    dd = np.linalg.norm(sites - ck, axis=1)
                    free[dd < rin[i] + rin_min] = False
                    free[k] = False
                    placed = True
            w0 += window
        if not placed:
            raise ValueError(f"deposition could not place particle {i}")
        while start < sites.shape[0] and not free[start]:
            start += 1
    return centers


def grid_for(box: np.ndarray, max_particlesize: float):
    """Cell grid of MWDiem.py:575-580 and the key range of coms inside the box."""
    xmax, xmin, ymax, ymin, zmax, zmin = box
    ncx = int((xmax - xmin) / max_particlesize)
    ncy = int((ymax - ymin) / max_particlesize)
    ncz = int((zmax - zmin) / max_particlesize)
    cell = np.array([(xmax - xmin) / ncx, (ymax - ymin) / ncy, (zmax - zmin) / ncz])
    lo = [int(xmin / cell[0]), int(ymin / cell[1]), int(zmin / cell[2])]
    hi = [int(xmax / cell[0]), int(ymax / cell[1]), int(zmax / cell[2])]
    kmin = lo[0] + lo[1] * ncx + lo[2] * ncx * ncy
    kmax = hi[0] + hi[1] * ncx + hi[2] * ncx * ncy
    return cell, ncx, ncy, ncz, kmin, kmax - kmin + 1


def make_scene(n_particles: int, box_mm, seed: int, npts=(12, 12), axes=(1.0, 0.8, 0.6), size=(0.8, 1.2),
               density: float = RHO_KG_MM3, law: Law4 | None = None, wall_law: Law4 | None = None,
               dt: float = 1e-6, g=(0.0, 0.0, -G_MM), wall_thickness_factor: float = 10.0,
               spacing_factor: float = 1.05, ref_inertia: bool = False, name: str = "",
               site_fn=None, device_hulls: bool = False) -> SceneData:
    """A bed of `n_particles` hulls of npts[0]..npts[1] points on an ellipsoid
    with semi-axes axes*U(size) in a walled box (the configurations' generator).
    The cell size (max_particlesize) is 2x the largest circumradius.
    site_fn(rmax, rng) -> [n, 3], if given, places the particles instead of
    the jittered lattice. device_hulls: hull bodies built on the GPU
    (make_particles_device; the scene then has no face lists)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    law = law or Law4()
    wall_law = wall_law or law
    if device_hulls:
        if ref_inertia:
            raise ValueError("device_hulls: the reference's volume element is built on the host only")
        parts = make_particles_device(n_particles, seed, npts, axes, size, density,
                                      consistent=(law.mode == PHYSICAL))
    else:
        parts = make_particles(n_particles, seed, npts, axes, size, density, ref_inertia,
                               consistent=(law.mode == PHYSICAL))
    rmax = max(float(np.max(np.linalg.norm(p[0], axis=1))) for p in parts) if parts else 1.0
    ms = 2.0 * rmax
    spacing = spacing_factor * ms
    if site_fn is not None:
        sites = site_fn(rmax, rng)
    else:
        try:
            sites = bcc_sites(box_mm, spacing, rmax, n_particles)
            jit = (spacing - ms) / 2.0 / math.sqrt(3.0)
            sites = sites + rng.uniform(-jit, jit, size=sites.shape)
        except ValueError:
            # too tight for bounding spheres: deposit with an exact separation test
            if any(p[5] is None for p in parts):
                raise ValueError("the lattice is too tight and device-built hulls carry no face normals "
                                 "for the deposit; build the hulls on the host")
            sites = deposit_sites([p[0] for p in parts], [p[5] for p in parts], box_mm, rng)
    T = wall_thickness_factor * ms
    walls = walls_for_box(box_mm, T, density, ref_inertia)
    W = len(walls)
    N = W + n_particles
    vmax = max([w[0].shape[0] for w in walls] + [p[0].shape[0] for p in parts])
    com = np.zeros((N, 3))
    verts = np.zeros((N, vmax, 3))
    nverts = np.zeros(N, dtype=np.int32)
    mass = np.zeros(N)
    I = np.zeros((N, 3, 3))
    Iinv = np.zeros((N, 3, 3))
    isize = np.zeros(N)
    wall_n = np.zeros((N, 3))
    wall_c = np.zeros((N, 3))
    faces = []
    for k, (V, n, c, m, Ibf, Iiv, isz, qf) in enumerate(walls):
        faces.append(qf)
        com[k] = c
        verts[k, :V.shape[0]] = V
        nverts[k] = V.shape[0]
        mass[k], I[k], Iinv[k], isize[k] = m, Ibf, Iiv, isz
        wall_n[k], wall_c[k] = n, c
    for i, (V, m, Ibf, Iiv, isz, _, qf) in enumerate(parts):
        faces.append(qf)
        r = W + i
        com[r] = sites[i]
        verts[r, :V.shape[0]] = V + sites[i]
        nverts[r] = V.shape[0]
        mass[r], I[r], Iinv[r], isize[r] = m, Ibf, Iiv, isz
    allv = np.concatenate([w[0] for w in walls])
    box = np.array([allv[:, 0].max(), allv[:, 0].min(), allv[:, 1].max(), allv[:, 1].min(),
                    allv[:, 2].max(), allv[:, 2].min()])
    cell, ncx, ncy, ncz, kmin, kspan = grid_for(box, ms)
    fext = mass[:, None] * np.asarray(g)[None, :]
    fext[:W] = 0.0
    return SceneData(n_walls=W, vmax=vmax, com=com, verts=verts, nverts=nverts, mass=mass, I=I, Iinv=Iinv,
                     isize=isize, fext=fext, text=np.zeros((N, 3)), wall_n=wall_n, wall_c=wall_c, box=box,
                     cell=cell, ncx=ncx, ncy=ncy, ncz=ncz, kmin=kmin, kspan=kspan, max_particlesize=ms,
                     law_p=law, law_w=wall_law, dt=dt, name=name, inner=tuple(float(x) for x in box_mm),
                     faces=None if device_hulls else faces)


# ------------------------------------------------------- the configurations
def config_small(seed: int = 1) -> SceneData:
    """200 hulls of 12 points, 10x10x30 mm box (plain method, small scene)."""
    return make_scene(200, (10.0, 10.0, 30.0), seed, npts=(12, 12), name="C1-small-200")


def config_large(seed: int = 2, n: int = 100_000) -> SceneData:
    """100,000 hulls of 8-20 points, 200x200x400 mm box (plain method, large settle)."""
    return make_scene(n, (200.0, 200.0, 400.0), seed, npts=(8, 20), name=f"C2-large-{n}")


def config_pile(seed: int = 5, n: int = 4096, v_pour: float = 1000.0, law: Law4 | None = None):
    """The differentiable form's pile: n hulls of 8-20 points on ellipsoids of
    semi-axes (1.0, 0.8, 0.6) x U(0.6, 0.9) mm, packed from the floor of a
    40x40x60 mm box (bounding spheres 2% apart; 4096 of the larger C2-sized
    hulls would not fit as a lattice) and poured, i.e. given random velocities
    of rms v_pour mm/s (1 m/s: a fall of about 5 cm), so contacts open and
    close through the run. Physical
    law Kn=1e4, en=et=0.5, mu=0.3. Returns the scene and the initial
    velocities [n_rows, 3] (zero for walls)."""
    law = law or Law4(Kn=1e4, en=0.5, et=0.5, mu=0.3, mode=PHYSICAL)
    sc = make_scene(n, (40.0, 40.0, 60.0), seed, npts=(8, 20), size=(0.6, 0.9), law=law, spacing_factor=1.02,
                    name=f"C5-pile-{n}")
    rng = np.random.Generator(np.random.PCG64([seed, 11]))
    v = np.zeros((sc.n_rows, 3))
    v[sc.n_walls:] = rng.normal(0.0, v_pour / math.sqrt(3.0), size=(n, 3))
    return sc, v


@dataclass
class Spawn:
    """A spawn schedule: rows [n_walls, n_active) take part from the start;
    at the start of every `every`-th step (the first included) `per` more
    rows join, in row order (MWDiem.py:685-714 creates particles at the start
    of every 5th step)."""
    n_active: int
    every: int
    per: int

    def active_at(self, step: int, n_rows: int) -> int:
        """Rows [0, n) active during step `step` (counted from 0)."""
        return min(n_rows, self.n_active + (step // self.every + 1) * self.per)


def config_pour(seed: int = 5, n: int = 4096, L=(40.0, 40.0, 60.0), v_pour: float = 2000.0, every: int | None = None,
                law: Law4 | None = None, npts=(8, 20), size=(0.8, 1.2), dt: float = 1e-6):
    """A pile formed by pouring: n hulls (the configurations' generator: 8-20
    points on ellipsoids of semi-axes (1.0, 0.8, 0.6) x U(0.8, 1.2) mm) enter
    the top of an L mm box in batches of one layer -- a square grid of sites
    2.1 x the largest circumradius apart, 1.5 circumradii under the ceiling,
    random orientation, a small jitter -- moving down at v_pour mm/s (2 m/s by
    default); a batch every `every` steps (by default the steps the previous
    batch takes to fall 1.05 grid spacings, clear of the layer). Physical law
    Kn=1e4, en=et=0.5, mu=0.3. Returns (scene, initial velocities [n_rows, 3],
    Spawn schedule): rows wait in place until their batch spawns."""
    law = law or Law4(Kn=1e4, en=0.5, et=0.5, mu=0.3, mode=PHYSICAL)
    Lx, Ly, Lz = L
    batch = {}

    def sites(rmax, rng):
        a = 2.1 * rmax
        xs = np.arange(rmax * 1.1, Lx - rmax * 1.1 + 1e-9, a)
        ys = np.arange(rmax * 1.1, Ly - rmax * 1.1 + 1e-9, a)
        X, Y = np.meshgrid(xs, ys, indexing="ij")
        layer = np.stack([X.ravel(), Y.ravel(), np.full(X.size, Lz - 1.5 * rmax)], axis=1)
        B = layer.shape[0]
        batch["per"] = B
        out = np.concatenate([layer[rng.permutation(B)] for _ in range((n + B - 1) // B)])[:n]
        jit = 0.04 * rmax
        return out + rng.uniform(-jit, jit, size=out.shape)

    sc = make_scene(n, L, seed, npts=npts, size=size, law=law, dt=dt, site_fn=sites, name=f"pour-{n}")
    rng = np.random.Generator(np.random.PCG64([seed, 12]))
    v = np.zeros((sc.n_rows, 3))
    v[sc.n_walls:] = rng.normal(0.0, 0.02 * v_pour, size=(n, 3))
    v[sc.n_walls:, 2] -= v_pour
    if every is None:
        every = int(math.ceil(1.05 * 2.1 * (sc.max_particlesize / 2) / (v_pour * dt)))
    # a packed pile of these hulls holds about 100 candidate pairs per hull
    # on cells twice the largest circumradius wide
    sc.pairs_per_particle = 200.0
    return sc, v, Spawn(n_active=sc.n_walls, every=every, per=batch["per"])


def compact_dense(sc: SceneData, gap: float = -0.02, seed: int = 0) -> SceneData:
    """Copy of a scene with the particles moved onto a BCC lattice of spacing
    (1+gap)*2*r_max from the floor up, so neighbouring hulls overlap (a
    contact-heavy state for stage tests). Only centres and vertices change."""
    out = sc.copy()
    W = sc.n_walls
    rmax = sc.max_particlesize / 2
    sites = bcc_sites(sc.inner, (1.0 + gap) * 2 * rmax, rmax, sc.n_particles)
    rng = np.random.Generator(np.random.PCG64(seed))
    sites = sites + rng.uniform(-0.01 * rmax, 0.01 * rmax, size=sites.shape)
    for i in range(sc.n_particles):
        r = W + i
        d = sites[i] - out.com[r]
        out.com[r] = sites[i]
        out.verts[r, :out.nverts[r]] += d
    return out


# ------------------------------------------------------------ batched form
def shape_pool(n_shapes: int, seed: int, npts=(8, 20), axes=(1.0, 0.8, 0.6), size=(0.8, 1.2),
               density: float = RHO_KG_MM3):
    """n_shapes distinct hulls (body frame, centred on the vertex mean) with
    mass, inertia size and inertia tensor; scenes draw from it and give each
    particle its own random orientation."""
    rng = np.random.Generator(np.random.PCG64([seed, 7]))
    out = []
    for _ in range(n_shapes):
        k = int(rng.integers(npts[0], npts[1] + 1))
        sz = float(rng.uniform(size[0], size[1]))
        pts = ellipsoid_points(rng, k, np.asarray(axes) * sz)
        V, tri, vol = hull_body(pts)
        V = V - V.mean(axis=0)
        I, isize = polyhedron_inertia(V, tri, density)
        out.append((V, density * vol, isize, I))
    return out


def random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - w * z)
    R[:, 0, 2] = 2 * (x * z + w * y)
    R[:, 1, 0] = 2 * (x * y + w * z)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - w * x)
    R[:, 2, 0] = 2 * (x * z - w * y)
    R[:, 2, 1] = 2 * (y * z + w * x)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def hopper_scene(n_particles: int, seed: int, pool, law: Law4, orifice: float = 12.0, bin_height: float = 40.0,
                 density: float = RHO_KG_MM3, dt: float = 1e-6, name: str = "") -> SceneData:
    """A flat-bottomed silo discharging through a slot orifice: particles seeded
    on a jittered lattice in the upper chamber, whose floor is two plates with
    a gap of `orifice` mm; a collection bin of `bin_height` mm below. Particles
    come from a shape pool with a random orientation each. Walls: floor,
    ceiling, x-, x+, y-, y+ and the two orifice plates (8 walls)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    idx = rng.integers(0, len(pool), size=n_particles)
    prad = np.array([float(np.max(np.linalg.norm(p[0], axis=1))) for p in pool])
    rmax = float(prad[idx].max())
    ms = 2.0 * rmax
    spacing = 1.05 * ms
    dens = math.sqrt(2.0) / spacing ** 3          # BCC sites per mm^3
    w = max(4 * ms + orifice, (n_particles / dens / 1.5) ** (1.0 / 3.0) + 2 * ms)
    h_fill = 1.5 * w
    for _ in range(20):
        try:
            sites = bcc_sites((w, w, h_fill), spacing, rmax, n_particles)
            break
        except ValueError:
            h_fill *= 1.15
    z0 = bin_height
    Lz = z0 + h_fill + ms
    sites = sites + np.array([0.0, 0.0, z0])
    jit = (spacing - ms) / 2.0 / math.sqrt(3.0)
    sites = sites + rng.uniform(-jit, jit, size=sites.shape)
    T = 10.0 * ms
    walls = walls_for_box((w, w, Lz), T, density, False)
    tp = 2.0 * ms  # orifice plate thickness
    for lo, hi in (((0.0, 0.0, z0 - tp), (w / 2 - orifice / 2, w, z0)),
                   ((w / 2 + orifice / 2, 0.0, z0 - tp), (w, w, z0))):
        P = box_corners(lo, hi)
        V, tri, vol = hull_body(P)
        c = V.mean(axis=0)
        I, isize = polyhedron_inertia(V - c, tri, density)
        walls.append((V, np.array([0.0, 0.0, 1.0]), c, density * vol, I, np.linalg.inv(I), isize))
    W = len(walls)
    N = W + n_particles
    vmax = max([wl[0].shape[0] for wl in walls] + [p[0].shape[0] for p in pool])
    com = np.zeros((N, 3))
    verts = np.zeros((N, vmax, 3))
    nverts = np.zeros(N, dtype=np.int32)
    mass = np.zeros(N)
    I = np.zeros((N, 3, 3))
    Iinv = np.zeros((N, 3, 3))
    isize = np.zeros(N)
    wall_n = np.zeros((N, 3))
    wall_c = np.zeros((N, 3))
    for k, (V, nrm, c, m, Ik, Iik, isz, *_) in enumerate(walls):
        com[k] = c
        verts[k, :V.shape[0]] = V
        nverts[k] = V.shape[0]
        mass[k], I[k], Iinv[k], isize[k] = m, Ik, Iik, isz
        wall_n[k], wall_c[k] = nrm, c
    # vectorised: pooled body vertices and tensors, one random rotation each
    pv = max(p[0].shape[0] for p in pool)
    Vp = np.zeros((len(pool), pv, 3))
    npool = np.array([p[0].shape[0] for p in pool], dtype=np.int32)
    for j, p in enumerate(pool):
        Vp[j, :p[0].shape[0]] = p[0]
    Ip = np.stack([p[3] for p in pool])
    R = random_rotations(rng, n_particles)
    Vw = np.einsum("nij,nkj->nki", R, Vp[idx])
    Iw = np.einsum("nij,njk,nlk->nil", R, Ip[idx], R)
    pad = np.arange(pv)[None, :] >= npool[idx][:, None]
    Vw = Vw + sites[:, None, :]
    Vw[pad] = 0.0
    com[W:] = sites
    verts[W:, :pv] = Vw
    nverts[W:] = npool[idx]
    mass[W:] = np.array([p[1] for p in pool])[idx]
    I[W:] = Iw
    Iinv[W:] = np.linalg.inv(Iw)
    isize[W:] = np.array([p[2] for p in pool])[idx]
    allv = np.concatenate([wl[0] for wl in walls])
    box = np.array([allv[:, 0].max(), allv[:, 0].min(), allv[:, 1].max(), allv[:, 1].min(),
                    allv[:, 2].max(), allv[:, 2].min()])
    cell, ncx, ncy, ncz, kmin, kspan = grid_for(box, ms)
    fext = mass[:, None] * np.array([0.0, 0.0, -G_MM])[None, :]
    fext[:W] = 0.0
    return SceneData(n_walls=W, vmax=vmax, com=com, verts=verts, nverts=nverts, mass=mass, I=I, Iinv=Iinv,
                     isize=isize, fext=fext, text=np.zeros((N, 3)), wall_n=wall_n, wall_c=wall_c, box=box,
                     cell=cell, ncx=ncx, ncy=ncy, ncz=ncz, kmin=kmin, kspan=kspan, max_particlesize=ms,
                     law_p=law, law_w=law, dt=dt, name=name, inner=(w, w, Lz))


def config_polydisperse(seed: int = 3, n: int = 1_000_000, L: float = 400.0, r_range=(0.5, 2.5), npts=(6, 32),
                        v_top: float = 100.0, n_shapes: int = 4096, law: Law4 | None = None,
                        density: float = RHO_KG_MM3, dt: float = 1e-6) -> SceneData:
    """The polydisperse stress test: n hulls (the convex hull of 6-32 points on
    an ellipsoid with semi-axes 1.0/0.8/0.6, shapes from a pool, each scaled
    to a circumradius log-uniform over r_range mm, random orientation) in an
    L^3 mm box of six walls whose ceiling moves along +x at v_top mm/s, under
    gravity, on the physical law.

    Placement: n must be a cube m^3; the hulls sit on an m^3 cubic lattice
    filling the box, the larger half of the radii on the sites of one parity
    and the smaller half on the other, so that two lattice neighbours never
    have meeting bounding spheres (checked; ValueError otherwise). The cell
    size is twice the largest circumradius."""
    m = int(round(n ** (1.0 / 3.0)))
    if m ** 3 != n:
        raise ValueError("n must be a cube (the lattice has m^3 sites)")
    law = law or Law4(mode=PHYSICAL)
    rng = np.random.Generator(np.random.PCG64(seed))
    pool = shape_pool(n_shapes, seed, npts=npts, size=(1.0, 1.0), density=density)
    prad = np.array([float(np.max(np.linalg.norm(p[0], axis=1))) for p in pool])
    idx = rng.integers(0, len(pool), size=n)
    r = np.exp(rng.uniform(math.log(r_range[0]), math.log(r_range[1]), size=n))
    order = np.argsort(-r, kind="stable")            # largest first
    g = np.arange(m, dtype=np.float64)
    margin = r_range[1] + 0.1
    spacing = (L - 2.0 * margin) / (m - 1) if m > 1 else 0.0
    ii, jj, kk = np.meshgrid(g, g, g, indexing="ij")
    par = ((ii + jj + kk).astype(np.int64) % 2).ravel()
    sites = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1) * spacing + margin
    even, odd = np.flatnonzero(par == 0), np.flatnonzero(par == 1)
    big, small = order[:even.size], order[even.size:]
    if m > 1 and r[big].max() + r[small].max() >= spacing:
        raise ValueError("bounding spheres of lattice neighbours would meet; enlarge L or narrow r_range")
    if m > 1 and 2.0 * r.max() >= spacing * math.sqrt(2.0):
        raise ValueError("bounding spheres of same-parity neighbours would meet")
    # rows in lattice order (neighbouring rows are neighbours in space, as a
    # generator that fills a box layer by layer numbers them): the radius of
    # row i is the one its site's parity received
    r_site = np.empty(n)
    r_site[even] = r[big]
    r_site[odd] = r[small]
    r = r_site
    site_of = np.arange(n, dtype=np.int64)
    scale = r / prad[idx]
    ms = 2.0 * float(r.max())
    T = 10.0 * ms
    walls = walls_for_box((L, L, L), T, density, False)
    W = len(walls)
    N = W + n
    pv = max(p[0].shape[0] for p in pool)
    vmax = max([wl[0].shape[0] for wl in walls] + [pv])
    com = np.zeros((N, 3))
    verts = np.zeros((N, vmax, 3))
    nverts = np.zeros(N, dtype=np.int32)
    mass = np.zeros(N)
    I = np.zeros((N, 3, 3))
    Iinv = np.zeros((N, 3, 3))
    isize = np.zeros(N)
    wall_n = np.zeros((N, 3))
    wall_c = np.zeros((N, 3))
    for k, (V, nrm, c, mw, Ik, Iik, isz, *_) in enumerate(walls):
        com[k] = c
        verts[k, :V.shape[0]] = V
        nverts[k] = V.shape[0]
        mass[k], I[k], Iinv[k], isize[k] = mw, Ik, Iik, isz
        wall_n[k], wall_c[k] = nrm, c
    Vp = np.zeros((len(pool), pv, 3))
    npool = np.array([p[0].shape[0] for p in pool], dtype=np.int32)
    for j, p in enumerate(pool):
        Vp[j, :p[0].shape[0]] = p[0]
    Ip = np.stack([p[3] for p in pool])
    mp = np.array([p[1] for p in pool])
    ip = np.array([p[2] for p in pool])
    pos = sites[site_of]
    # in chunks: rotated, scaled body vertices and tensors (mass ~ s^3, inertia ~ s^5)
    for a in range(0, n, 1 << 16):
        b = min(n, a + (1 << 16))
        R = random_rotations(rng, b - a)
        s3 = scale[a:b]
        Vw = np.einsum("nij,nkj->nki", R, Vp[idx[a:b]]) * s3[:, None, None]
        Iw = np.einsum("nij,njk,nlk->nil", R, Ip[idx[a:b]], R) * (s3 ** 5)[:, None, None]
        pad = np.arange(pv)[None, :] >= npool[idx[a:b]][:, None]
        Vw = Vw + pos[a:b, None, :]
        Vw[pad] = 0.0
        rows = slice(W + a, W + b)
        verts[rows, :pv] = Vw
        I[rows] = Iw
        Iinv[rows] = np.linalg.inv(Iw)
    com[W:] = pos
    nverts[W:] = npool[idx]
    mass[W:] = mp[idx] * scale ** 3
    isize[W:] = ip[idx] * scale ** 5
    allv = np.concatenate([wl[0] for wl in walls])
    box = np.array([allv[:, 0].max(), allv[:, 0].min(), allv[:, 1].max(), allv[:, 1].min(),
                    allv[:, 2].max(), allv[:, 2].min()])
    cell, ncx, ncy, ncz, kmin, kspan = grid_for(box, ms)
    fext = mass[:, None] * np.array([0.0, 0.0, -G_MM])[None, :]
    fext[:W] = 0.0
    wall_v = np.zeros((N, 3))
    wall_v[1] = (v_top, 0.0, 0.0)  # the ceiling (walls_for_box order: floor, ceiling, x-, x+, y-, y+)
    return SceneData(n_walls=W, vmax=vmax, com=com, verts=verts, nverts=nverts, mass=mass, I=I, Iinv=Iinv,
                     isize=isize, fext=fext, text=np.zeros((N, 3)), wall_n=wall_n, wall_c=wall_c, box=box,
                     cell=cell, ncx=ncx, ncy=ncy, ncz=ncz, kmin=kmin, kspan=kspan, max_particlesize=ms,
                     law_p=law, law_w=law, dt=dt, name=f"polydisperse-{n}", inner=(L, L, L), wall_v=wall_v)


def config_batched(n_scenes: int = 256, seed: int = 4, n_min: int = 500, n_max: int = 20_000,
                   n_shapes: int = 4096, mode: int = PHYSICAL):
    """The batched form: n_scenes hopper-discharge scenes with particle counts
    log-uniform in [n_min, n_max], friction U(0.1,0.6), Kn log-uniform
    5e3-5e4, restitution U(0.2,0.8); scene s is drawn from seed 4*10^6+s."""
    pool = shape_pool(n_shapes, seed)
    scenes = []
    for s in range(n_scenes):
        rng = np.random.Generator(np.random.PCG64(4_000_000 * (seed // 4 if seed >= 4 else 1) + s))
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        mu = float(rng.uniform(0.1, 0.6))
        Kn = float(math.exp(rng.uniform(math.log(5e3), math.log(5e4))))
        e = float(rng.uniform(0.2, 0.8))
        law = Law4(Kn=Kn, en=e, et=e, mu=mu, mode=mode)
        scenes.append(hopper_scene(n, int(rng.integers(1 << 31)), pool, law, name=f"hopper-{s}-{n}"))
    return scenes

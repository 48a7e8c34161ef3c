"""Writes tests/golden/writer_golden.npz: the reference's particle generator
and output writers run on a few hulls, as input/output vectors for the
generator's hull construction and the VTK / particle-data writers.

This script imports the reference's pure-Python modules from its source tree
(CLASSES/Particle_class.py, LIBRARIES/LIBRARY_Vtk_creator.py,
LIBRARIES/LIBRARY_Write_particle_data.py), so it runs only where that tree is
present (by default /root/reference; --ref to change):

    python tests/golden/make_writer_golden.py [--ref PATH]

What it records, for 6 hulls made by Particle.Spherical_convex_polygon_maker_3D
(8-20 points on a sphere of radius N(1.0, 0.1) mm, numpy seed 20261204):
  * the points each hull was built from (the first ConvexHull input of each
    hull, captured by wrapping the module's ConvexHull name), and what the
    class made of them: vertices (centred), face vertex ids, volume, mass,
    inertia tensor and its size;
  * the hulls moved to fixed sites with fixed velocities, the VTK file
    Vtk_creator writes for rows [0, 6) (identity rotation, no pending
    translation) and the vertex coordinates it wrote;
  * the Data_<id>.dat files Particle_data_writer writes at timesteps 0 and 40
    (centres and velocities changed in between).
"""
from __future__ import annotations

import argparse
import os
import sys
import tempfile
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent / "writer_golden.npz"
N = 6
DENSITY = 2.5e-6  # kg/mm^3


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    args = ap.parse_args()
    ref = Path(args.ref)
    sys.path[:0] = [str(ref / "CLASSES"), str(ref / "LIBRARIES")]
    import LIBRARY_Vtk_creator as vtk_mod  # noqa: E402
    import LIBRARY_Write_particle_data as dat_mod  # noqa: E402
    import Particle_class as pc  # noqa: E402

    seen = []
    real_hull = pc.ConvexHull

    def recording_hull(points, *a, **k):
        seen.append(np.array(points, dtype=np.float64, copy=True))
        return real_hull(points, *a, **k)

    pc.ConvexHull = recording_hull
    np.random.seed(20261204)
    parts, pts = [], []
    for i in range(N):
        p = pc.Particle(i)
        before = len(seen)
        p.Spherical_convex_polygon_maker_3D(8, 21, 1.0, 0.1, DENSITY, 1, np.array([0.0, 0.0, -9810.0]),
                                            np.array([0.0, 0.0, 0.0]))
        pts.append(seen[before])
        parts.append(p)
    pc.ConvexHull = real_hull

    vmax = max(p.num_vertices for p in parts)
    fmax = max(p.num_faces for p in parts)
    pmax = max(a.shape[0] for a in pts)
    d = {"n": np.int64(N), "density": np.float64(DENSITY)}
    d["points"] = np.zeros((N, pmax, 3))
    d["npoints"] = np.array([a.shape[0] for a in pts], np.int32)
    d["verts0"] = np.zeros((N, vmax, 3))
    d["nverts"] = np.array([p.num_vertices for p in parts], np.int32)
    d["faces"] = -np.ones((N, fmax, 3), np.int32)
    d["nfaces"] = np.array([p.num_faces for p in parts], np.int32)
    d["volume"] = np.array([p.volume for p in parts])
    d["mass"] = np.array([p.mass for p in parts])
    d["inertia"] = np.array([p.moment_of_inertia for p in parts])
    d["inertia_size"] = np.array([p.moment_of_inertia_size for p in parts])
    for i, p in enumerate(parts):
        d["points"][i, :pts[i].shape[0]] = pts[i]
        for j in range(p.num_vertices):
            d["verts0"][i, j] = p.vertices[j].vertex_coo
        for j in range(p.num_faces):
            d["faces"][i, j] = p.faces[j].face_vertices_id

    rng = np.random.Generator(np.random.PCG64(7))
    sites = rng.uniform(2.0, 8.0, size=(N, 3))
    vel = rng.normal(0.0, 300.0, size=(N, 3))
    for i, p in enumerate(parts):
        p.Translate(sites[i])
        p.velocity = vel[i].copy()
    with tempfile.TemporaryDirectory() as tmp:
        cwd = os.getcwd()
        os.chdir(tmp)
        try:
            qall = np.tile(np.array([1.0, 0.0, 0.0, 0.0]), (N, 1))
            dxall = np.zeros((N, 3))
            vtk_mod.Vtk_creator(parts, "Out_0.vtk", 0, N, qall, dxall)
            d["vtk"] = np.frombuffer(Path("Out_0.vtk").read_bytes(), np.uint8)
            d["verts_written"] = np.zeros((N, vmax, 3))
            d["com_written"] = np.array([np.asarray(p.center_of_mass, float) for p in parts])
            for i, p in enumerate(parts):
                for j in range(p.num_vertices):
                    d["verts_written"][i, j] = p.vertices[j].vertex_coo
            d["com0"] = d["com_written"].copy()
            d["vel0"] = vel.copy()
            for p in parts:
                dat_mod.Particle_data_writer(p, "Data_", 0)
            com1 = d["com0"] + rng.normal(0.0, 0.5, size=(N, 3))
            vel1 = rng.normal(0.0, 300.0, size=(N, 3))
            for i, p in enumerate(parts):
                p.center_of_mass = com1[i].copy()
                p.velocity = vel1[i].copy()
                dat_mod.Particle_data_writer(p, "Data_", 40)
            d["com1"], d["vel1"] = com1, vel1
            for i in range(N):
                d[f"dat_{i}"] = np.frombuffer(Path(f"Data_{i}.dat").read_bytes(), np.uint8)
        finally:
            os.chdir(cwd)
    np.savez_compressed(OUT, **d)
    print(f"wrote {OUT}: {N} hulls, vertices {list(d['nverts'])}, faces {list(d['nfaces'])}")


if __name__ == "__main__":
    main()

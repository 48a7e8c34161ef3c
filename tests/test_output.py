"""Output formats (legacy-VTK POLYDATA snapshots, Data_<id>.dat series):
layout of LIBRARY_Vtk_creator.py:3-41 and LIBRARY_Write_particle_data.py:3-12."""
from __future__ import annotations

import numpy as np
import pytest

from mwdiem_amd import output


def _cube(shift=(0.0, 0.0, 0.0), vmax=10):
    c = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], float) + shift
    V = np.zeros((vmax, 3))
    V[:8] = c
    return V


def test_hull_faces_are_outward_triangles():
    verts = np.stack([_cube(), _cube((3.0, 0.5, 0.25))])
    nv = np.array([8, 8], np.int32)
    faces = output.hull_faces(verts, nv)
    for r in range(2):
        f = faces[r]
        assert f.shape == (12, 3) and f.min() >= 0 and f.max() < 8
        V = verts[r, :8]
        cm = V.mean(0)
        n = np.cross(V[f[:, 1]] - V[f[:, 0]], V[f[:, 2]] - V[f[:, 1]])
        assert np.all(np.einsum("ij,ij->i", n, V[f[:, 0]] - cm) > 0)


def test_vtk_layout_and_round_trip(tmp_path):
    rng = np.random.Generator(np.random.PCG64(3))
    verts = np.stack([_cube(), _cube((2.0, 0.0, 0.0)), _cube((0.0, 2.0, 0.0))])
    verts[:, :8] += rng.normal(scale=1e-3, size=(3, 8, 3))
    nv = np.array([8, 8, 8], np.int32)
    faces = output.hull_faces(verts, nv)
    p = tmp_path / "Out_1.vtk"
    output.write_vtk(p, verts, nv, faces, 1, 3)
    lines = p.read_text().split("\n")
    assert lines[:4] == ["# vtk DataFile Version 2.0", "# kata", "ASCII", "DATASET POLYDATA"]
    assert lines[4] == "POINTS\t16\tfloat"
    pts = np.array([[float(x) for x in ln.split("\t")] for ln in lines[5:21]])
    assert np.array_equal(pts, np.concatenate([verts[1, :8], verts[2, :8]]))  # exact round trip
    assert lines[21] == f"POLYGONS\t{len(faces[1]) + len(faces[2])}\t{4 * (len(faces[1]) + len(faces[2]))}"
    tri = np.array([[int(x) for x in ln.split("\t")] for ln in lines[22:-1]])
    assert np.all(tri[:, 0] == 3)
    assert np.array_equal(tri[:12, 1:], faces[1]) and np.array_equal(tri[12:, 1:], faces[2] + 8)
    assert lines[-1] == ""


def test_data_lines_match_the_reference_formatting(tmp_path):
    rng = np.random.Generator(np.random.PCG64(4))
    com = rng.normal(size=(3, 3)) * 10
    v = rng.normal(size=(3, 3))
    prefix = str(tmp_path / "Data_")
    output.write_particle_data(prefix, [1, 2], 0, com, v)
    output.write_particle_data(prefix, [1, 2], 1000, com * 2, v)
    txt = (tmp_path / "Data_2.dat").read_text().split("\n")
    assert txt[0] + "\n" == output.DATA_HEADER
    # the reference writes str() of numpy float64 values, |vz|, and integer zeros
    want = "\t".join(["0"] + [str(np.float64(x)) for x in com[2]] + [str(np.float64(v[2, 0])),
                      str(np.float64(v[2, 1])), str(abs(np.float64(v[2, 2])))] + ["0"] * 9)
    assert txt[1] == want
    assert txt[2].startswith("1000\t") and txt[3] == ""
    output.write_particle_data(prefix, [2], 0, com, v)  # timestep 0 starts the file again
    assert len((tmp_path / "Data_2.dat").read_text().split("\n")) == 3


@pytest.mark.gpu
def test_async_writer_matches_synchronous_files(gpu, dense_phys, tmp_path):
    """The writer's snapshots are taken back to back with the steps (no
    synchronisation between a snapshot and the next step); the same steps on
    a second stepper, synchronised and written directly, give the same
    files."""
    from mwdiem_amd.stepper import Stepper
    S = Stepper(dense_phys)
    Wr = output.AsyncWriter(S, tmp_path / "async")
    for step in range(3):
        Wr.vtk(step)
        Wr.data(step)
        S.step(5)
    Wr.close()
    R = Stepper(dense_phys)
    ref_dir = tmp_path / "sync"
    ref_dir.mkdir()
    assert dense_phys.faces is not None  # the generator's faces, which the writer prints
    faces = [np.asarray(f, dtype=np.int64) for f in dense_phys.faces]
    W = dense_phys.n_walls
    for step in range(3):
        st = R.numpy_state()
        output.write_vtk(ref_dir / f"Out_{step}.vtk", st["verts"], dense_phys.nverts, faces, 0, dense_phys.n_rows)
        output.write_particle_data(str(ref_dir / "Data_"), range(W, dense_phys.n_rows), step, st["com"], st["v"])
        R.step(5)
    for step in range(3):
        assert (tmp_path / "async" / f"Out_{step}.vtk").read_text() == (ref_dir / f"Out_{step}.vtk").read_text()
    for r in range(W, dense_phys.n_rows):
        assert (tmp_path / "async" / f"Data_{r}.dat").read_text() == (ref_dir / f"Data_{r}.dat").read_text()
    # the snapshots differ from step to step (the steps moved the bed)
    assert (tmp_path / "async" / "Out_0.vtk").read_text() != (tmp_path / "async" / "Out_2.vtk").read_text()

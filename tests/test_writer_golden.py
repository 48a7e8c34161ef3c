"""The generator's hull construction and the output writers against the
reference's own code (tests/golden/writer_golden.npz, written by
tests/golden/make_writer_golden.py, which runs
CLASSES/Particle_class.py:258-303 and the writers
LIBRARIES/LIBRARY_Vtk_creator.py:3-41 and LIBRARY_Write_particle_data.py:3-12).

* scene.hull_body on the points each reference hull was built from gives its
  vertices (in Qhull's vertex order, centred on their mean), its faces (Qhull's
  facet order, each facet's vertex order -- what Vtk_creator prints), its
  volume; polyhedron_inertia with the reference's volume element gives its
  inertia tensor and size.
* output.vtk_text on the vertex coordinates Vtk_creator wrote, with those
  faces, is the reference's file byte for byte; write_particle_data writes
  the reference's Data_<id>.dat byte for byte over two timesteps.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from mwdiem_amd import output, scene

FIX = Path(__file__).resolve().parent / "golden" / "writer_golden.npz"


def _fix():
    return dict(np.load(FIX))


def test_hull_construction_matches_reference_generator():
    d = _fix()
    rho = float(d["density"])
    for i in range(int(d["n"])):
        P = d["points"][i, :d["npoints"][i]]
        V, tri, vol, qtri = scene.hull_body(P, qhull_faces=True)
        nv, nf = int(d["nverts"][i]), int(d["nfaces"][i])
        assert V.shape[0] == nv and qtri.shape[0] == nf
        cm = np.zeros(3)
        for v in V:  # the reference's running sum (Particle_class.py:221-224)
            cm = cm + v
        cm = cm / nv
        assert np.abs((V - cm) - d["verts0"][i, :nv]).max() <= 1e-15
        assert np.array_equal(qtri, d["faces"][i, :nf])
        assert abs(vol - d["volume"][i]) <= 1e-15 * d["volume"][i]
        assert abs(rho * vol - d["mass"][i]) <= 1e-15 * d["mass"][i]
        I, isize = scene.polyhedron_inertia(V - cm, tri, rho, reference_volume_element=True)
        assert np.abs(I - d["inertia"][i]).max() <= 1e-12 * np.abs(d["inertia"][i]).max()
        assert abs(isize - d["inertia_size"][i]) <= 1e-12 * d["inertia_size"][i]


def test_vtk_writer_is_byte_identical():
    d = _fix()
    n = int(d["n"])
    faces = [scene.hull_body(d["points"][i, :d["npoints"][i]], qhull_faces=True)[3] for i in range(n)]
    text = output.vtk_text(d["verts_written"], d["nverts"], [f.astype(np.int64) for f in faces], 0, n)
    assert text.encode() == bytes(d["vtk"])


def test_particle_data_writer_is_byte_identical(tmp_path):
    d = _fix()
    n = int(d["n"])
    prefix = str(tmp_path / "Data_")
    output.write_particle_data(prefix, range(n), 0, d["com0"], d["vel0"])
    output.write_particle_data(prefix, range(n), 40, d["com1"], d["vel1"])
    for i in range(n):
        assert Path(f"{prefix}{i}.dat").read_bytes() == bytes(d[f"dat_{i}"]), i

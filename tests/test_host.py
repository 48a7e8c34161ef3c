"""Host-side checks that need no GPU: the C ABI exports, the drop-in
signatures, the scene generator."""
from __future__ import annotations

import inspect
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    txt = (ROOT / "include" / "mwdiem_hip.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mwd_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import ctypes
    from mwdiem_amd import _lib
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    names = _declared()
    assert len(names) >= 19
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes table binds exactly the declared entry points
    assert sorted(_lib.EXPORTED) == names


def test_law_init_without_gpu():
    import ctypes
    from mwdiem_amd import _lib
    L = _lib.law(1e4, 0.5, 0.5, 0.3, 1e-6)
    # tc = sqrt(mu (pi^2 + ln^2 en) / Kn), Gn = -2 mu ln(en) / tc (.cu:690-691)
    ln = np.log(0.5)
    tc = np.sqrt(0.3 * (np.pi ** 2 + ln * ln) / 1e4)
    assert abs(L.tc - tc) <= 1e-15 * tc and abs(L.Gn - (-2 * 0.3 * ln / tc)) <= 1e-12 * abs(L.Gn)
    with pytest.raises(_lib.MwdError):
        _lib.law(-1.0, 0.5, 0.5, 0.3, 1e-6)
    assert ctypes.sizeof(_lib.Law) == 14 * 8


REF_SIGS = {
    # LIBRARIES/gpu_fine_contact_detection.py:76-89
    "Contact_detection_GPU": "particle_A_center_of_mass particle_B_center_of_mass particle_A_vertices "
                             "particle_B_vertices particle_A_velocity particle_B_velocity particle_A_angular_velocity "
                             "particle_B_angular_velocity particle_A_numvertices particle_B_numvertices particle_A_id "
                             "particle_B_id displacement_tang Kn en et mu Kn_wall en_wall et_wall mu_wall "
                             "num_particles_wall force torque num_vertices_maxi quaternion_A quaternion_B "
                             "timestep_size mom_inertia_size_A mom_inertia_size_B",
    # LIBRARIES/gpu_Verlet.py:53-61
    "Verlet_GPU": "velocity_arr velocity_12_arr angular_velocity_12_bf_arr quaternion_12_arr quaternion_arr "
                  "quaternion_all_arr angular_velocity_bf_arr angular_velocity_arr force_arr torque_arr vertices_arr "
                  "center_of_mass_arr timestep_size mass_arr num_particles moment_of_inertia_inv_arr "
                  "moment_of_inertia_arr dx_all_arr num_vertices_max syssize_x_max syssize_x_min syssize_y_max "
                  "syssize_y_min syssize_z_max syssize_z_min num_part_min num_part_max",
    # LIBRARIES/LIBRARY_Coarse_contact_detection.py:84
    "Linked_cell_hash": "particle numcells_x numcells_y num_part_min num_part_max pairstocheck hashh_arr",
}


def test_dropin_signatures_match_reference():
    from mwdiem_amd.dropin import LIBRARY_Coarse_contact_detection as CC
    from mwdiem_amd.dropin import gpu_fine_contact_detection as FC
    from mwdiem_amd.dropin import gpu_Verlet as GV
    for mod, name in ((FC, "Contact_detection_GPU"), (GV, "Verlet_GPU"), (CC, "Linked_cell_hash")):
        params = list(inspect.signature(getattr(mod, name)).parameters)
        assert params == REF_SIGS[name].split(), name


def test_dropins_fail_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from mwdiem_amd.dropin import gpu_Verlet as GV
    with pytest.raises(RuntimeError):
        GV.Verlet_GPU(*([np.zeros((2, 3))] * 12), 1e-6, np.ones((2, 1)), 2, np.zeros((2, 3, 3)), np.zeros((2, 3, 3)),
                      np.zeros((2, 3)), 1, 1, -1, 1, -1, 1, -1, 0, 2)


def test_scene_generator_invariants():
    from mwdiem_amd import scene
    sc = scene.make_scene(40, (8.0, 8.0, 20.0), 5)
    W = sc.n_walls
    assert W == 6 and sc.n_particles == 40
    # centre of mass = vertex mean (Particle_class.py:264-268)
    for r in range(W, sc.n_rows):
        assert np.allclose(sc.verts[r, :sc.nverts[r]].mean(axis=0), sc.com[r], atol=1e-12)
    # every particle inside the inner box; every centre's key inside the key range
    inner = np.array(sc.inner)
    for r in range(W, sc.n_rows):
        v = sc.verts[r, :sc.nverts[r]]
        assert np.all(v >= 0) and np.all(v <= inner)
    from oracle import oracle as O
    keys = O.cell_keys(sc.com[W:], sc.cell, sc.ncx, sc.ncy)
    assert keys.min() >= sc.kmin and keys.max() < sc.kmin + sc.kspan
    # the cell is 2x the largest circumradius (about the vertex mean)
    rmax = max(np.max(np.linalg.norm(sc.verts[r, :sc.nverts[r]] - sc.com[r], axis=1)) for r in range(W, sc.n_rows))
    assert abs(sc.max_particlesize - 2 * rmax) < 1e-12
    # no particle overlaps another at t = 0 (oracle GJK on all pairs)
    for a in range(W, sc.n_rows):
        for b in range(a + 1, sc.n_rows):
            if np.linalg.norm(sc.com[a] - sc.com[b]) < sc.max_particlesize:
                c, *_ = O.gjk_epa(sc.verts[a, :sc.nverts[a]], sc.verts[b, :sc.nverts[b]], sc.com[a], sc.com[b])
                assert c != 1, (a, b)


def test_inertia_of_a_cube():
    # a cube of side s and mass m about its centre: I = m s^2 / 6 (identity)
    from mwdiem_amd import scene
    s, rho = 2.0, 3.0
    P = scene.box_corners((-1, -1, -1), (1, 1, 1))
    V, tri, vol = scene.hull_body(P)
    I, isize = scene.polyhedron_inertia(V, tri, rho)
    m = rho * vol
    assert abs(vol - 8.0) < 1e-12
    assert np.allclose(I, np.eye(3) * m * s * s / 6, rtol=1e-12, atol=1e-12)
    assert abs(isize - m * s * s / 6) < 1e-12


def test_law_tangent_matches_central_differences():
    """mwd_law_tangent's derivatives of every coefficient against central
    differences of mwd_law_init_mode, for a seed that mixes the inputs."""
    from mwdiem_amd import _lib
    base = np.array([1e4, 0.5, 0.4, 0.3])
    seed = np.array([[1.0, 0.0, 2.0], [0.0, 1.0, 0.5], [0.0, 1.0, 0.0], [0.3, 0.0, 1.0]])  # d input / d theta
    T = _lib.law_tangent(*base, seed)
    names = ("Kn", "mu", "Gn", "Kt_num", "inv_mu", "ln_et", "cn", "ct", "ratio")
    for k in range(3):
        h = 1e-6
        xp = base + h * seed[:, k]
        xm = base - h * seed[:, k]
        Lp = _lib.law(*xp, 1e-6, _lib.LAW_PHYSICAL)
        Lm = _lib.law(*xm, 1e-6, _lib.LAW_PHYSICAL)
        for n in names:
            fd = (getattr(Lp, n) - getattr(Lm, n)) / (2 * h)
            got = getattr(T, n)[k]
            assert abs(got - fd) <= 1e-6 * max(1.0, abs(fd)), (n, k, got, fd)


def test_cell_scale_tuning_refuses_reference_law(small_scene):
    """The reference law's candidate order depends on its grid, so the cell
    size is not tuned there (refused before any GPU work)."""
    from mwdiem_amd.tune import regrid, tune_cell_scale
    with pytest.raises(ValueError):
        tune_cell_scale([small_scene])
    wide = regrid(small_scene, 1.5)
    assert wide.ncx < small_scene.ncx and np.all(wide.cell >= 1.5 * small_scene.max_particlesize)
    assert wide.kspan < small_scene.kspan

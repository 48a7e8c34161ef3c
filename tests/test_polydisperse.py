"""The polydisperse stress configuration (generator, moving ceiling) and the
FP32 contact reduction: CPU checks of the generator and the oracle's wall
motion; GPU checks against the oracle and between the reduction precisions,
and the full 1,000,000-hull configuration."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O


def _radii(sc):
    W = sc.n_walls
    live = np.arange(sc.vmax)[None, :] < sc.nverts[W:, None]
    d = np.linalg.norm(sc.verts[W:] - sc.com[W:, None, :], axis=2)
    d[~live] = 0.0
    return d.max(axis=1)


def test_generator_sizes_and_placement():
    from scipy.spatial import cKDTree
    from mwdiem_amd import scene
    sc = scene.config_polydisperse(n=8000, L=80.0)
    W = sc.n_walls
    r = _radii(sc)
    assert r.min() >= 0.5 * (1 - 1e-9) and r.max() <= 2.5 * (1 + 1e-9)
    assert abs(np.median(r) / np.sqrt(0.5 * 2.5) - 1.0) < 0.05      # log-uniform: median sqrt(a b)
    assert sc.nverts[W:].min() >= 4 and sc.nverts[W:].max() <= 32
    pairs = cKDTree(sc.com[W:]).query_pairs(6.0, output_type="ndarray")
    d = np.linalg.norm(sc.com[W:][pairs[:, 0]] - sc.com[W:][pairs[:, 1]], axis=1)
    assert len(pairs) > 0 and np.all(d > r[pairs[:, 0]] + r[pairs[:, 1]])  # no bounding spheres meet
    assert sc.max_particlesize == pytest.approx(2 * r.max())
    assert sc.moving_walls and np.array_equal(sc.wall_v[1], [100.0, 0.0, 0.0])
    assert np.count_nonzero(sc.wall_v[:W].any(axis=1)) == 1
    # mass and inertia scale as s^3 and s^5 of the pooled shape: density is uniform
    tr = np.trace(sc.I[W:], axis1=1, axis2=2)
    assert np.all(tr > 0) and np.allclose(sc.isize[W:], tr / 3.0, rtol=1e-9)


def test_oracle_moves_walls_by_v_dt():
    from mwdiem_amd import scene
    sc = scene.config_polydisperse(n=1000, L=40.0)
    st = O.initial_state(sc)
    c0, v0, wc0 = st["com"].copy(), st["verts"].copy(), st["wall_c"].copy()
    for _ in range(7):
        O.step(st, sc)
    W = sc.n_walls
    d = 7 * sc.dt * 100.0
    assert np.allclose(st["com"][1] - c0[1], [d, 0, 0], atol=1e-12)
    assert np.allclose(st["wall_c"][1] - wc0[1], [d, 0, 0], atol=1e-12)
    assert np.allclose(st["verts"][1, :8] - v0[1, :8], [d, 0, 0], atol=1e-12)
    assert np.array_equal(st["com"][[0, 2, 3, 4, 5]], c0[[0, 2, 3, 4, 5]])


def _agitated(sc, seed=3, vrms=3000.0, vz=2000.0):
    """Random velocities (mm/s) with an upward drift: hulls meet each other
    and the ceiling within a few hundred steps."""
    st = O.initial_state(sc)
    W = sc.n_walls
    rng = np.random.Generator(np.random.PCG64(seed))
    v = rng.normal(size=(sc.n_particles, 3)) * vrms
    v[:, 2] += vz
    st["v"][W:] = v
    st["v12"][W:] = v
    return st


@pytest.mark.gpu
def test_gpu_moving_ceiling_matches_oracle(gpu):
    """1,000 agitated hulls thrown up into the ceiling, which slides along x:
    after 250 steps positions and quaternions within 1e-8 of the oracle, and
    the walls where the oracle puts them."""
    from mwdiem_amd import scene
    from mwdiem_amd.stepper import Stepper
    sc = scene.config_polydisperse(n=1000, L=40.0)
    st = _agitated(sc)
    S = Stepper(sc)
    S.set_state({k: st[k] for k in ("v", "v12")})
    n_c = 0
    for _ in range(250):
        _, coll = O.step(st, sc)
        n_c += int((coll == 1).sum())
        S.step(1)
    S.check()
    got = S.numpy_state()
    W = sc.n_walls
    assert n_c > 50
    for k in ("com", "q", "v", "w"):
        b = st[k][W:]
        assert np.abs(got[k][W:] - b).max() <= 1e-8 * np.abs(b).max(), k
    assert np.abs(got["com"][:W] - st["com"][:W]).max() <= 1e-12
    assert np.abs(S.t["wall_c"].cpu().numpy()[:W] - st["wall_c"][:W]).max() <= 1e-12


@pytest.mark.gpu
def test_gpu_fp32_reduction_close_to_fp64(gpu):
    """The single-precision contact sums differ from the double ones (the
    path is taken) but stay within 1e-4 of the displacement after 250 steps
    of agitated hulls."""
    from mwdiem_amd import scene
    from mwdiem_amd.stepper import Stepper
    sc = scene.config_polydisperse(n=1000, L=40.0)
    st = _agitated(sc)
    runs = {}
    for fp32 in (False, True):
        S = Stepper(sc, reduce_fp32=fp32)
        S.set_state({k: st[k] for k in ("v", "v12")})
        S.step(250)
        S.check()
        runs[fp32] = S.numpy_state()
    W = sc.n_walls
    a, b = runs[True]["com"][W:], runs[False]["com"][W:]
    disp = np.abs(b - sc.com[W:]).max()
    diff = np.abs(a - b).max()
    print(f"fp32 vs fp64: max |dx| {diff:.3e} of displacement {disp:.3e}")
    assert diff > 0.0 and diff <= 1e-4 * disp


@pytest.mark.gpu
def test_gpu_full_polydisperse_configuration(gpu):
    """1,000,000 hulls, radii 0.5-2.5 mm, 6-32 points, 400 mm box, sliding
    ceiling: the first candidate list equals the oracle's, order included;
    then the hulls are thrown about (5 m/s rms, so lattice neighbours meet
    within the run) for 300 steps in FP64 and in FP32 reductions: both runs
    see thousands of contacts, the FP32 sums change the result (the path is
    taken), the median hull moves by less than 1e-9 of the displacement
    between the two and fewer than 0.1 % of the hulls part by more than
    1 um."""
    import time
    from mwdiem_amd import scene
    from mwdiem_amd.stepper import Stepper
    t0 = time.time()
    sc = scene.config_polydisperse()
    print(f"generated in {time.time() - t0:.1f} s", flush=True)
    S = Stepper(sc)
    S.candidates()
    s0 = S.check()
    ref = O.candidates(O.initial_state(sc), sc)
    got = S.pairs()
    print(f"pairs {got.shape[0]} (oracle {ref.shape[0]}), cells {s0['n_cells']}", flush=True)
    assert got.shape == ref.shape and np.array_equal(got, ref)
    del ref, got
    S.close()
    W = sc.n_walls
    rng = np.random.Generator(np.random.PCG64(3))
    v = np.zeros((sc.n_rows, 3))
    v[W:] = rng.normal(size=(sc.n_particles, 3)) * 5000.0
    v[:W] = sc.wall_v[:W]
    out = {}
    for fp32 in (False, True):
        S = Stepper(sc, reduce_fp32=fp32)
        S.set_state({"v": v, "v12": v})
        S.step(300)
        out[fp32] = (S.numpy_state()["com"], S.check())
        S.close()
        print(f"fp32={fp32}: {out[fp32][1]}", flush=True)
    a, b = out[True][0][W:], out[False][0][W:]
    disp = np.abs(b - sc.com[W:]).max()
    dx = np.linalg.norm(a - b, axis=1)
    q = np.quantile(dx, [0.5, 0.99, 0.999])
    print(f"fp32 vs fp64 |dx|: median {q[0]:.3e}, 99% {q[1]:.3e}, 99.9% {q[2]:.3e}, max {dx.max():.3e}, "
          f"above 1 um {np.mean(dx > 1e-3):.2e}; displacement {disp:.3e}")
    assert out[False][1]["n_contacts"] > 1000 and out[True][1]["n_contacts"] > 1000
    # the FP32 sums change the result (the path is taken); a colliding hull
    # gas is chaotic, so the two runs are compared by the share of hulls
    # that have parted company (a few collide violently enough to diverge)
    assert dx.max() > 0.0
    assert q[0] <= 1e-9 * disp and np.mean(dx > 1e-3) < 1e-3

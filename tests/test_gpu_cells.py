"""GPU parity of the cell classification and fresh-cell kernels against the oracle, through
the C ABI: labels, ghost indices, nearest triangles and points bit-exact; fresh-cell values
bit-exact."""
import numpy as np
import pytest

from conftest import random_small_triangles

pytestmark = pytest.mark.gpu

KEYS = ("type", "ghost_idx", "nearest", "point", "reflect", "dist")


def _same_cells(a, b):
    for k in KEYS:
        x, y = a[k], b[k]
        if x.dtype == np.float32:
            x, y = x.view(np.uint32), y.view(np.uint32)
        assert np.array_equal(x, y), (k, int(np.sum(x != y)))


@pytest.mark.parametrize("thin_shell", [False, True])
def test_sphere_classification_bitexact(amd, oracle_mod, scenes_mod, thin_shell):
    sc = scenes_mod.c1_static_sphere(steps=2)
    print(f"classify: icosphere L3 on {sc.res}^3, thin_shell={thin_shell}", flush=True)
    g = amd.PPPMSolver(sc.res, sc.dl, sc.dt, sc.min_pos, 0, 0.83, sc.c)
    o = oracle_mod.Oracle(sc.res, sc.dl, sc.dt, sc.min_pos, sc.c, 0.83, oracle_mod.FP64)
    g.set_mesh(sc.verts, sc.tris)
    o.set_mesh(sc.verts, sc.tris)
    assert g.classify(thin_shell) == o.classify(thin_shell)
    _same_cells(g.cells(), o.cells())


def test_random_open_triangles_classification_bitexact(amd, oracle_mod):
    """Scattered open triangles (unknown cells, column fill, ties in the nearest search)."""
    rng = np.random.default_rng(31)
    N, h = 24, 0.005
    dt = h / (np.sqrt(3) * 343.0)
    verts, tris = random_small_triangles(rng, 80, N, h, size=1.5, margin=4)
    print(f"classify: 80 random triangles on {N}^3", flush=True)
    g = amd.PPPMSolver(N, h, dt, (0, 0, 0), 0, 0.83, 343.0)
    o = oracle_mod.Oracle(N, h, dt, (0, 0, 0), 343.0, 0.83, oracle_mod.FP64)
    g.set_mesh(verts, tris)
    o.set_mesh(verts, tris)
    assert g.classify() == o.classify()
    _same_cells(g.cells(), o.cells())


def test_moving_scene_classification(amd, oracle_mod, scenes_mod):
    """Moving sphere + cube on 64^3: labels after several mesh updates agree bit for bit."""
    sc = scenes_mod.c2_moving(steps=40)
    print(f"classify: moving scene {sc.res}^3, T={sc.tris.shape[0]}", flush=True)
    g = amd.PPPMSolver(sc.res, sc.dl, sc.dt, sc.min_pos, 0, 0.83, sc.c)
    o = oracle_mod.Oracle(sc.res, sc.dl, sc.dt, sc.min_pos, sc.c, 0.83, oracle_mod.FP64)
    g.set_mesh(sc.motion(0), sc.tris)
    o.set_mesh(sc.motion(0), sc.tris)
    assert g.classify() == o.classify()
    _same_cells(g.cells(), o.cells())
    for i in (13, 27, 39):
        v = sc.motion(i)
        g.update_mesh(v)
        o.update_mesh(v)
        assert g.classify() == o.classify()
        _same_cells(g.cells(), o.cells())


def test_fresh_cells_of_a_translating_sphere(amd, oracle_mod, scenes_mod):
    """Sphere on 32^3 translated by 0.45 h per update, 6 steps between updates: the fresh-cell
    fill of the stepped grid (levels t and t-1) and set_solid_zero agree bit for bit."""
    sc = scenes_mod.c1_static_sphere(steps=60)
    print(f"fresh cells: translating sphere on {sc.res}^3", flush=True)
    g = amd.PPPMSolver(sc.res, sc.dl, sc.dt, sc.min_pos, 0, 0.83, sc.c)
    o = oracle_mod.Oracle(sc.res, sc.dl, sc.dt, sc.min_pos, sc.c, 0.83, oracle_mod.FP64)
    g.set_mesh(sc.verts, sc.tris)
    o.set_mesh(sc.verts, sc.tris)
    assert g.classify() == o.classify()
    rng = np.random.default_rng(32)
    total = 0
    v = sc.verts.copy()
    for u in range(5):
        for i in range(6):
            nm = sc.neumann[6 * u + i]
            g.update_grid_and_face(nm)
            o.step(nm)
        v[:, 0] += np.float32(0.45 * sc.dl)
        g.update_mesh(v)
        o.update_mesh(v)
        assert g.classify() == o.classify()
        _same_cells(g.cells(), o.cells())
        a_now = rng.standard_normal(sc.tris.shape[0]).astype(np.float32)
        a_old = rng.standard_normal(sc.tris.shape[0]).astype(np.float32)
        # the fill reads the two grid levels: start both sides from the oracle's
        for rel in (0, -1):
            g.set_grid(rel, o.grid(rel))
        nf = g.fill_fresh(a_now, a_old)
        assert nf == o.fill_fresh(a_now, a_old)
        total += max(nf, 0)
        for rel in (0, -1):
            assert np.array_equal(g.get_grid(rel).view(np.uint32), o.grid(rel).view(np.uint32)), (u, rel)
    assert total > 0
    g.set_solid_zero()
    o.set_solid_zero()
    assert np.array_equal(g.get_grid(0).view(np.uint32), o.grid(0).view(np.uint32))
    print("fresh cells filled:", total, flush=True)

"""Pins of the oracle's cell classification and fresh cells (ghost_cell_util.cu:7-230,
ghost_cell.cu:51-218, helper_math.h:1603-1800) against independent geometry: a float64
point-triangle distance, the true inside/outside of a sphere, the ghost-cell definition, and
the closed form of the fresh-cell fill for a uniform field."""
import math

import numpy as np
import pytest

AIR, SOLID, GHOST, UNKNOWN = 0, 1, 2, 3


def _seg_dist(p, a, b):
    ab = b - a
    t = np.clip(np.dot(p - a, ab) / np.dot(ab, ab), 0.0, 1.0)
    return np.linalg.norm(p - (a + t * ab))


def _tri_dist(p, v0, v1, v2):
    """float64 distance from p to a triangle: the plane projection if it falls inside,
    otherwise the nearest of the three edges."""
    n = np.cross(v1 - v0, v2 - v0)
    n2 = np.dot(n, n)
    q = p - np.dot(p - v0, n) / n2 * n
    # barycentrics of the projection
    w0 = np.dot(np.cross(v1 - q, v2 - q), n) / n2
    w1 = np.dot(np.cross(v2 - q, v0 - q), n) / n2
    w2 = 1 - w0 - w1
    if w0 >= 0 and w1 >= 0 and w2 >= 0:
        return np.linalg.norm(p - q)
    return min(_seg_dist(p, v0, v1), _seg_dist(p, v1, v2), _seg_dist(p, v2, v0))


def test_closest_point_matches_independent_distance(oracle_mod):
    rng = np.random.default_rng(21)
    for _ in range(3000):
        tri = rng.standard_normal((3, 3)).astype(np.float32)
        p = (rng.standard_normal(3) * rng.choice([0.2, 1.0, 5.0])).astype(np.float32)
        w = oracle_mod.closest_coeff(p, tri)
        assert np.all(w >= -1e-5) and abs(float(w.sum()) - 1.0) < 1e-5
        T = tri.astype(np.float64)
        q = w[0] * T[0] + w[1] * T[1] + w[2] * T[2]
        d = np.linalg.norm(p.astype(np.float64) - q)
        ref = _tri_dist(p.astype(np.float64), *T)
        assert abs(d - ref) <= 1e-5 * max(1.0, ref), (d, ref)


@pytest.fixture(scope="module")
def sphere_cells(oracle_mod, scenes_mod):
    sc = scenes_mod.c1_static_sphere(steps=2)
    o = oracle_mod.Oracle(sc.res, sc.dl, sc.dt, sc.min_pos, sc.c, 0.83, oracle_mod.FP64)
    o.set_mesh(sc.verts, sc.tris)
    g = o.classify()
    return sc, o, g, o.cells()


def test_sphere_inside_outside(sphere_cells):
    sc, o, g, cells = sphere_cells
    N, h = sc.res, sc.dl
    idx = np.arange(N ** 3)
    xyz = np.stack([idx // (N * N), idx // N % N, idx % N], 1)
    ctr = (xyz + 0.5) * h + np.array(sc.min_pos)
    centre = sc.verts.astype(np.float64).mean(0)
    radius = np.linalg.norm(sc.verts.astype(np.float64) - centre, axis=1).mean()
    d = np.linalg.norm(ctr - centre, axis=1) - radius
    t = cells["type"]
    assert np.all(np.isin(t[d < -0.3 * h], [SOLID, GHOST]))
    assert np.all(t[d > 0.3 * h] == AIR)
    assert not np.any(t == UNKNOWN)


def test_ghost_cells_are_solid_cells_next_to_air(sphere_cells):
    sc, o, g, cells = sphere_cells
    N = sc.res
    t = cells["type"].reshape(N, N, N)
    solidish = (t == SOLID) | (t == GHOST)
    air = t == AIR
    nb_air = np.zeros_like(air)
    for ax in range(3):
        for sh in (1, -1):
            rolled = np.roll(air, sh, axis=ax)
            sl = [slice(None)] * 3
            sl[ax] = 0 if sh == 1 else N - 1
            rolled[tuple(sl)] = False
            nb_air |= rolled
    assert np.array_equal(t == GHOST, solidish & nb_air)
    gi = cells["ghost_idx"].reshape(N, N, N)
    # ghost indices enumerate the ghost cells in linear order
    assert np.array_equal(np.sort(gi[t == GHOST].ravel()), np.arange(g))
    assert np.array_equal(gi[t == GHOST].ravel(), np.arange(g))
    assert np.all(gi[t != GHOST] == -1)


def test_nearest_distance_is_the_mesh_distance(sphere_cells):
    sc, o, g, cells = sphere_cells
    N, h = sc.res, sc.dl
    V = sc.verts.astype(np.float64)
    T = sc.tris
    rng = np.random.default_rng(22)
    near = np.where(cells["dist"] < 1.5 * h)[0]
    for c in rng.choice(near, 150, replace=False):
        x, y, z = c // (N * N), c // N % N, c % N
        p = (np.array([x, y, z]) + 0.5) * h + np.array(sc.min_pos)
        ref = min(_tri_dist(p, *V[tri]) for tri in T)
        assert abs(cells["dist"][c] - ref) <= 1e-5 * h, (c, cells["dist"][c], ref)
        # the stored point is on the nearest triangle and the mirror point is 2 q - p
        q = cells["point"][c].astype(np.float64)
        assert abs(_tri_dist(q, *V[T[cells["nearest"][c]]])) <= 1e-6 * h
        assert np.allclose(cells["reflect"][c], 2 * q - p, atol=1e-6 * h)


def test_fresh_cells_of_a_uniform_field(oracle_mod, scenes_mod):
    """A uniform field C around a sphere that moves by 0.6 h: every fresh cell (not AIR before,
    AIR after) gets C - rho * a * |xf - xr| at t and t-1 when an AIR cell lies in the 2x2x2
    block around its mirror point (normalised inverse-distance weights of a constant), and the
    nearest such cell's C otherwise (ghost_cell.cu:75-196)."""
    sc = scenes_mod.c1_static_sphere(steps=2)
    N, h = sc.res, sc.dl
    o = oracle_mod.Oracle(N, h, sc.dt, sc.min_pos, sc.c, 0.83, oracle_mod.FP64)
    o.set_mesh(sc.verts, sc.tris)
    o.classify()
    old = o.cells()["type"].copy()
    shift = np.float32(0.6 * h)
    moved = sc.verts.copy()
    moved[:, 0] += shift
    o.update_mesh(moved)
    o.classify()
    cells = o.cells()
    C = np.float32(0.75)
    o.set_grid(0, np.full(N ** 3, C, np.float32))
    o.set_grid(-1, np.full(N ** 3, C, np.float32))
    rng = np.random.default_rng(23)
    a_now = rng.standard_normal(len(sc.tris)).astype(np.float32)
    a_old = rng.standard_normal(len(sc.tris)).astype(np.float32)
    n = o.fill_fresh(a_now, a_old)
    fresh = np.where((old != AIR) & (cells["type"] == AIR))[0]
    assert n == len(fresh) > 0
    g0, g1 = o.grid(0).ravel(), o.grid(-1).ravel()
    both_air = (old == AIR) & (cells["type"] == AIR)
    for c in fresh:
        x, y, z = c // (N * N), c // N % N, c % N
        xf = (np.array([x, y, z]) + 0.5) * h + np.array(sc.min_pos)
        xb = cells["point"][c].astype(np.float64)
        xr = xb + 2 * (xf - xb)
        b = np.floor((xr - np.array(sc.min_pos)) / h - 0.5).astype(int)
        block = [(b[0] + i, b[1] + j, b[2] + k) for i in (0, 1) for j in (0, 1) for k in (0, 1)]
        if any(both_air[(p[0] * N + p[1]) * N + p[2]] for p in block):
            L = np.linalg.norm(xf - xr)
            tri = cells["nearest"][c]
            assert g0[c] == pytest.approx(C - 1.225 * a_now[tri] * L, rel=1e-5, abs=1e-6)
            assert g1[c] == pytest.approx(C - 1.225 * a_old[tri] * L, rel=1e-5, abs=1e-6)
        else:
            assert g0[c] == C and g1[c] == C
    # cells that are not fresh keep C
    other = np.ones(N ** 3, bool)
    other[fresh] = False
    assert np.all(g0[other] == C) and np.all(g1[other] == C)


def test_set_solid_zero(oracle_mod, scenes_mod):
    sc = scenes_mod.c1_static_sphere(steps=2)
    N = sc.res
    o = oracle_mod.Oracle(N, sc.dl, sc.dt, sc.min_pos, sc.c, 0.83, oracle_mod.FP64)
    o.set_mesh(sc.verts, sc.tris)
    o.classify()
    o.set_grid(0, np.ones(N ** 3, np.float32))
    o.set_solid_zero()
    t = o.cells()["type"]
    g = o.grid(0).ravel()
    assert np.all(g[t == SOLID] == 0) and np.all(g[t != SOLID] == 1)


def _stencil_scene(oracle_mod, air_cells, fresh=(4, 5, 5), xr_cells=(8.75, 5.5, 5.5)):
    """16^3 grid (h = 5 cm, origin 0), every cell SOLID before and after except the given cells,
    AIR before and after, and one fresh cell (SOLID before, AIR after) whose nearest surface
    point is placed so that its mirror point 2 xf - xb lands at xr_cells (in cell units)."""
    N, h = 16, 0.05
    o = oracle_mod.Oracle(N, h, h / (math.sqrt(3) * 343.0), (0, 0, 0), 343.0, 0.83, oracle_mod.FP64)
    tri = np.array([[12.1, 12.2, 12.3], [12.6, 12.2, 12.3], [12.1, 12.7, 12.3]], np.float32) * h
    o.set_mesh(tri, np.array([[0, 1, 2]], np.int32))
    NN = N ** 3
    lin = lambda p: (p[0] * N + p[1]) * N + p[2]  # noqa: E731
    now = np.full(NN, SOLID, np.int32)
    old = np.full(NN, SOLID, np.int32)
    for p in air_cells:
        now[lin(p)] = old[lin(p)] = AIR
    now[lin(fresh)] = AIR
    xf = (np.array(fresh) + 0.5) * h
    xb = 2 * xf - np.array(xr_cells) * h
    point = np.zeros((NN, 3), np.float32)
    point[lin(fresh)] = xb
    nearest = np.zeros(NN, np.int32)
    o.set_cells(now, old, nearest, point)
    rng = np.random.default_rng(31)
    g0, g1 = rng.standard_normal(NN).astype(np.float32), rng.standard_normal(NN).astype(np.float32)
    o.set_grid(0, g0)
    o.set_grid(-1, g1)
    return o, g0, g1, lin, np.linalg.norm(xf - np.array(xr_cells) * h)


def test_fresh_fill_weights_are_inverse_distance_squared(oracle_mod):
    """Shepard interpolation with power 2 (ghost_cell.cu:120-131): the mirror point sits a
    quarter of the way from the AIR cell (8,5,5) to the AIR cell (9,5,5), the other six cells
    of its 2x2x2 block are not AIR, so the weights are 1/(h/4)^2 : 1/(3h/4)^2 = 9 : 1
    (inverse distance would give 3 : 1), each with the reference's EPS = 1e-8 added to the
    squared distance. The acceleration term is -1.225 a |xf - xr|."""
    A, B = (8, 5, 5), (9, 5, 5)
    o, g0, g1, lin, L = _stencil_scene(oracle_mod, [A, B])
    acc_now, acc_old = np.float32([0.7]), np.float32([-0.4])
    assert o.fill_fresh(acc_now, acc_old) == 1
    c = lin((4, 5, 5))
    n0, n1 = o.grid(0).ravel(), o.grid(-1).ravel()
    h = 0.05
    wa, wb = 1 / ((h / 4) ** 2 + 1e-8), 1 / ((3 * h / 4) ** 2 + 1e-8)
    assert wa / wb == pytest.approx(9.0, rel=2e-4)
    for got, g, acc in ((n0[c], g0, acc_now[0]), (n1[c], g1, acc_old[0])):
        want = (wa * g[lin(A)] + wb * g[lin(B)]) / (wa + wb) - 1.225 * acc * L
        assert got == pytest.approx(want, rel=1e-5, abs=1e-6)
    # nothing else changes
    mask = np.ones(len(n0), bool)
    mask[c] = False
    assert np.array_equal(n0[mask], g0[mask]) and np.array_equal(n1[mask], g1[mask])


def test_fresh_fill_falls_back_to_the_nearest_air_cell(oracle_mod):
    """No AIR cell in the 2x2x2 block of the mirror point: the value (at t and t-1) is copied
    from the nearest cell that is AIR before and after within the 4x4x4 block around the base
    cell, without the acceleration term (ghost_cell.cu:136-160). Of (10,5,5) at 1.25 h and
    (7,6,7) at sqrt(2.5625 + 1 + 2.25) h, the first is nearer."""
    near, far = (10, 5, 5), (7, 6, 7)
    o, g0, g1, lin, _ = _stencil_scene(oracle_mod, [near, far])
    assert o.fill_fresh(np.float32([5.0]), np.float32([5.0])) == 1
    c = lin((4, 5, 5))
    assert o.grid(0).ravel()[c] == g0[lin(near)]
    assert o.grid(-1).ravel()[c] == g1[lin(near)]

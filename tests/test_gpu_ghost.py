"""GPU ghost-cell boundary solve (pppm_ghost_* through the C ABI) against the oracle on the
GPU's own classification: ghost list, order, condition numbers, stencil weights and sparse
rows after set-up; right-hand side and solution of one solve on a random pressure field; 40
steps of GhostCellSolver::update; the thin-shell (threshold 0) mode; a moving mesh.

Tolerances: the stencil weights come from an FP64 SVD on both sides (one-sided Jacobi here,
LAPACK in the oracle), so they agree to the FP32 rounding of FP64 values (relative 1e-6 is
asserted, differing FP32 codes are counted and printed). The solve runs BiCGSTAB in FP64 to
a relative residual of 1e-10, the oracle solves directly: relative 1e-5 is asserted."""
import numpy as np
import pytest

import ghost_oracle as go

pytestmark = pytest.mark.gpu


def _setup_pair(amd, sc, threshold, thin=False):
    g = amd.PPPMSolver(sc.res, sc.dl, sc.dt, sc.min_pos, 0, 0.83, sc.c)
    g.set_mesh(sc.verts, sc.tris)
    n = g.classify(thin)
    second = g.ghost_setup(threshold)
    cells = g.cells()
    st = go.ghost_setup(cells, sc.res, sc.dl, sc.min_pos, threshold)
    return g, n, second, cells, st


def _close(a, b, rel, what):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-30)
    err = np.abs(a - b).max() / scale
    print(f"{what}: max rel err {err:.3e}", flush=True)
    assert err <= rel, what


def test_ghost_setup_matches_oracle(amd, scenes_mod):
    sc = scenes_mod.c1_static_sphere(steps=2)
    print(f"ghost setup: icosphere L3 on {sc.res}^3", flush=True)
    g, n, second, cells, st = _setup_pair(amd, sc, 25.0)
    assert n == st["n"] > 100
    d = g.ghost_get(n)
    assert np.array_equal(d["cells"], st["cells"])
    assert np.array_equal(d["order"], st["order"])
    assert second == int(st["order"].sum()) > 0
    _close(d["cond"], st["cond"], 1e-5, "condition numbers")
    nbits = int(np.count_nonzero(d["weights"].view(np.uint32) != st["weights"].view(np.uint32)))
    print(f"weights: {nbits} of {d['weights'].size} FP32 codes differ", flush=True)
    _close(d["weights"], st["weights"], 1e-6, "weights")
    assert np.array_equal(d["cols"], st["cols"])
    _close(d["vals"], st["vals"], 1e-6, "matrix values")


@pytest.mark.parametrize("res", [32, 96])
def test_ghost_solve_matches_oracle(amd, scenes_mod, res):
    """res 32: the register/LDS solver (at most 3072 rows); res 96: the global-memory one."""
    sc = scenes_mod.c1_static_sphere(steps=2, res=res)
    g, n, _, cells, st = _setup_pair(amd, sc, 25.0)
    rng = np.random.default_rng(4)
    p = rng.standard_normal(sc.res ** 3).astype(np.float32)
    g.set_grid(0, p)
    neu = rng.standard_normal(sc.tris.shape[0]).astype(np.float32)
    g.ghost_set_neumann(neu)
    it = g.ghost_solve()
    print(f"ghost solve: {n} ghosts, {it} iterations", flush=True)
    assert 0 < it < 300
    d = g.ghost_get(n)
    b = go.ghost_rhs(st, cells, p, neu)
    _close(d["rhs"], b, 1e-5, "right-hand side")
    x = go.ghost_solve(st, b)
    _close(d["x"], x, 1e-5, "solution")
    grid = g.get_grid(0).reshape(-1)
    assert np.array_equal(grid[st["cells"]], d["x"])


def test_ghost_steps_match_oracle(amd, oracle_mod, scenes_mod):
    sc = scenes_mod.c1_static_sphere(steps=40)
    g, n, _, cells, st = _setup_pair(amd, sc, 25.0)
    o = oracle_mod.Oracle(sc.res, sc.dl, sc.dt, sc.min_pos, sc.c, 0.83, oracle_mod.FP64)
    for i in range(sc.steps):
        g.ghost_step(sc.neumann[i])
        go.ghost_step(o, st, cells, sc.neumann[i])
        if (i + 1) % 10 == 0:
            _close(g.get_grid(0), o.grid(0), 1e-5, f"grid after step {i + 1}")
            _close(g.get_grid(-1), o.grid(-1), 1e-5, f"previous level after step {i + 1}")


def test_thin_shell_first_order_bitexact(amd, oracle_mod, scenes_mod):
    """Threshold 0: first order everywhere, no linear solve: the ghost values are the
    right-hand side itself, formed in FP32 in the same order on both sides."""
    sc = scenes_mod.c1_static_sphere(steps=12)
    g, n, second, cells, st = _setup_pair(amd, sc, 0.0, thin=True)
    assert second == 0 and np.all(st["order"] == 0)
    o = oracle_mod.Oracle(sc.res, sc.dl, sc.dt, sc.min_pos, sc.c, 0.83, oracle_mod.FP64)
    for i in range(sc.steps):
        g.ghost_step(sc.neumann[i])
        go.ghost_step(o, st, cells, sc.neumann[i])
    a, b = g.get_grid(0), o.grid(0)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_ghost_update_mesh(amd, oracle_mod, scenes_mod):
    """GhostCellSolver::update_mesh on a translating sphere: the set-up after each update
    matches the oracle's on the GPU's classification."""
    sc = scenes_mod.c1_static_sphere(steps=6)
    g, n, _, cells, st = _setup_pair(amd, sc, 25.0)
    for k in range(1, 4):
        g.ghost_step(sc.neumann[k])
        v = sc.verts + np.float32(0.31 * k * sc.dl) * np.array([1, 0.5, 0], np.float32)
        fresh = g.ghost_update_mesh(v)
        cells = g.cells()
        st = go.ghost_setup(cells, sc.res, sc.dl, sc.min_pos, 25.0)
        d = g.ghost_get(st["n"])
        print(f"update {k}: {st['n']} ghosts, {fresh} fresh cells", flush=True)
        assert fresh >= 0
        assert np.array_equal(d["order"], st["order"])
        assert np.array_equal(d["cols"], st["cols"])
        _close(d["weights"], st["weights"], 1e-6, f"weights after update {k}")

"""Pins of the oracle's ghost-cell boundary solve (GhostCellSolver, src/algorithm/ghost_cell.cu:
220-545, svd.cu:129-148) on synthetic classifications of a planar wall:
  - the stencil weights satisfy Phi^T w = phi(r) (they interpolate every basis function, with
    the Neumann row at the nearest point), and reduce to trilinear weights without that row;
  - the condition number is the 2-norm condition number of Phi;
  - a linear pressure field with its normal derivative as Neumann data is reproduced exactly
    at the ghost cells, by the second-order system and by the first-order rule (stencil cells
    that are SOLID are read as they stand, so the field is given there too)."""
import numpy as np
import pytest

import ghost_oracle as go

F = np.float32


def plane_cells(N, h, mn, nrm, off):
    """Cells of the wall n.x < off (SOLID) in a grid: GHOST = SOLID with an AIR face neighbour,
    nearest point = the projection of the cell centre, mirror point = 2 nearest - centre."""
    nrm = np.asarray(nrm, np.float64)
    nrm = nrm / np.linalg.norm(nrm)
    ax = (np.arange(N).astype(F) + F(0.5)) * F(h)
    cen = np.stack(np.meshgrid(ax + F(mn[0]), ax + F(mn[1]), ax + F(mn[2]), indexing="ij"), -1).astype(F)
    s = cen.astype(np.float64) @ nrm - off
    typ = np.where(s < 0, go.SOLID, go.AIR).astype(np.int32)
    air = typ == go.AIR
    nb_air = np.zeros_like(air)
    for d in range(3):
        for sh in (1, -1):
            r = np.roll(air, sh, axis=d)
            sl = [slice(None)] * 3
            sl[d] = 0 if sh == 1 else N - 1
            r[tuple(sl)] = False
            nb_air |= r
    ghost = (typ == go.SOLID) & nb_air
    typ[ghost] = go.GHOST
    gidx = np.full(N ** 3, -1, np.int32)
    flat = ghost.reshape(-1)
    gidx[flat] = np.arange(int(flat.sum()), dtype=np.int32)
    near = (cen.astype(np.float64) - s[..., None] * nrm).astype(F)
    refl = (cen.astype(np.float64) - 2 * s[..., None] * nrm).astype(F)
    return {"type": typ.reshape(-1), "ghost_idx": gidx, "nearest": np.zeros(N ** 3, np.int32),
            "point": near.reshape(-1, 3), "reflect": refl.reshape(-1, 3), "dist": np.abs(s).astype(F).reshape(-1),
            "centers": cen.reshape(-1, 3), "normal": nrm}


N, H = 24, 0.01
MN = (-0.12, -0.12, -0.12)


@pytest.fixture(scope="module", params=[(1.0, 0.0, 0.0), (0.6, 0.7, 0.39), (-0.3, 0.2, 0.93)])
def wall(request):
    c = plane_cells(N, H, MN, request.param, 0.0031)
    return c, go.ghost_setup(c, N, H, MN, 25.0)


def test_weights_interpolate_the_basis(wall):
    """Second-order rows: sum_i w_i Phi[i, k] = phi_k(r) for every basis function k."""
    _, st = wall
    sec = st["order"] == 1
    assert sec.sum() > 10
    lhs = np.einsum("nik,ni->nk", st["phi"][sec].astype(np.float64), st["weights"][sec].astype(np.float64))
    np.testing.assert_allclose(lhs, st["phr"][sec].astype(np.float64), atol=2e-5)


def test_weights_without_neumann_row_are_trilinear(wall):
    c, st = wall
    sec = st["order"] == 1
    gc = np.stack([st["cells"] // (N * N), (st["cells"] // N) % N, st["cells"] % N], 1)
    self_in = np.zeros(st["n"], bool)
    for i in range(8):
        self_in |= np.all(st["base"] + go.CORNERS[i] == gc, axis=1)
    sel = sec & ~self_in
    if not sel.any():
        pytest.skip("every stencil of this wall contains its ghost")
    r = st["phr"][sel][:, 4:7].astype(np.float64)  # (x, y, z) of the mirror point in [-1, 1]^3
    tri = np.stack([np.prod([(1 + (2 * go.CORNERS[i][d] - 1) * r[:, d]) / 2 for d in range(3)], axis=0)
                    for i in range(8)], 1)
    np.testing.assert_allclose(st["weights"][sel], tri, atol=2e-6)


def test_condition_number_is_the_2norm_condition(wall):
    _, st = wall
    ref = np.linalg.cond(st["phi"].astype(np.float64))
    np.testing.assert_allclose(st["cond"], ref, rtol=1e-5)


def _linear_field(c, a, c0):
    p = (c["centers"].astype(np.float64) @ np.asarray(a) + c0).astype(F)
    return p


@pytest.mark.parametrize("nrm", [(1.0, 0.0, 0.0), (0.0, -1.0, 0.0)])
@pytest.mark.parametrize("threshold", [25.0, 0.0])
def test_axis_wall_reproduces_a_linear_field(nrm, threshold):
    """p = a.x + c0 on the AIR cells, a_n = a.n: the ghost values are a.x_g + c0."""
    c = plane_cells(N, H, MN, nrm, 0.0031)
    st = go.ghost_setup(c, N, H, MN, threshold)
    a = np.array([3.0, -2.0, 1.5])
    p = _linear_field(c, a, 0.7)
    p[c["type"] == go.GHOST] = 0  # unknowns
    neu = np.array([a @ np.asarray(nrm)], F)
    b = go.ghost_rhs(st, c, p, neu)
    x = go.ghost_solve(st, b)
    exact = (c["centers"][st["cells"]].astype(np.float64) @ a + 0.7)
    np.testing.assert_allclose(x, exact, atol=2e-5)
    if threshold == 0:
        assert np.all(st["order"] == 0) and np.array_equal(x, b)


def test_tilted_wall_second_order_rows_hold_for_the_exact_field(wall):
    """Row-local: with every ghost at its exact value, the second-order rows of A x = b hold."""
    c, st = wall
    a = np.array([1.0, 2.5, -0.5])
    p = _linear_field(c, a, -0.2)  # SOLID cells in a stencil are read as they are: the field too
    neu = np.array([a @ c["normal"]], F)
    b = go.ghost_rhs(st, c, p, neu).astype(np.float64)
    xe = c["centers"][st["cells"]].astype(np.float64) @ a - 0.2
    res = st["A"] @ xe - b
    sec = (st["order"] == 1) & st["inside"]  # a half-space wall meets the grid's faces
    assert sec.sum() > 10
    np.testing.assert_allclose(res[sec], 0.0, atol=5e-5)

"""CPU restatement of the reference's ghost-cell boundary solve — TEST INFRASTRUCTURE ONLY.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the checker.

    ghost_setup      construct_ghost_cell_list, construct_phi_matrix_kernel,
                     precompute_p_weight_kernel, construct_equation_kernel<true, false>,
                     COOMatrix::eliminate_zeros   src/algorithm/ghost_cell.cu:220-350, 352-495
                     cusolver_svd / inverse_from_svd_kernel   src/algorithm/svd.cu:36-157
    ghost_rhs        construct_equation_kernel<false, true>   ghost_cell.cu:352-470
    ghost_solve      BiCGSTAB_Solver::solve                    src/algorithm/sparse.cu:512-545
                     (here: the system's exact solution in FP64)
    ghost_step       GhostCellSolver::update                   src/algorithm/ghost_cell.h:131-141

FP32 quantities (stencil matrix, weights, right-hand side) are formed with numpy float32
element-wise operations in the reference's order; the SVD is numpy's FP64 LAPACK routine.
Cell data come from a classification (type, ghost index, nearest triangle, nearest point,
mirrored point, distance per cell, N^3 each, x-major).
"""
import numpy as np

AIR, SOLID, GHOST, UNKNOWN = 0, 1, 2, 3
F = np.float32


def _center(b, h, mn):
    """getCenter: ((float)i + 0.5f) * h + min, per component (particle_grid.h:172-175)."""
    return (b.astype(F) + F(0.5)) * F(h) + mn


def _normalized(p, b, h, mn):
    """get_normalized_coord (ghost_cell.cu:12-17): (p - centre) / h * 2 + (-1)."""
    d = (p - _center(b, h, mn)) / F(h)
    return d * F(2) + F(-1)


def _dot(a, b):
    return a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1] + a[..., 2] * b[..., 2]


def _base(refl, h, mn):
    """get_base_coord_for_reflect (ghost_cell.cu:20-26): int3((r - min) / h - 0.5f)."""
    q = (refl - mn) / F(h)
    return (q - F(0.5)).astype(np.int32)  # truncation toward zero, as the C cast


CORNERS = np.array([[i % 2, (i // 2) % 2, i // 4] for i in range(8)], np.int32)


def stencil_matrices(gc, near, refl, h, mn):
    """Phi (n, 8, 8) float32 (construct_phi_matrix_kernel, ghost_cell.cu:254-300), the base
    cell of each stencil, the unit normals and phi(reflect) (n, 8)."""
    n = gc.shape[0]
    dn = refl - near
    inv = F(1) / np.sqrt(_dot(dn, dn))
    normal = dn * inv[:, None]
    base = _base(refl, h, mn)
    phi = np.zeros((n, 8, 8), F)
    cn = _normalized(near, base, h, mn)
    zero = np.zeros(n, F)
    one = np.ones(n, F)
    for i in range(8):
        nb = base + CORNERS[i]
        self_ = np.all(nb == gc, axis=1)
        c = (CORNERS[i] * 2 - 1).astype(F)
        plain = np.array([c[0] * c[1] * c[2], c[0] * c[1], c[1] * c[2], c[0] * c[2], c[0], c[1], c[2], 1], F)
        phi[:, i, :] = plain
        if self_.any():
            cx, cy, cz = cn[:, 0], cn[:, 1], cn[:, 2]
            rows = [
                _dot(normal, np.stack([cy * cz, cx * cz, cx * cy], 1)),
                _dot(normal, np.stack([cy, cx, zero], 1)),
                _dot(normal, np.stack([zero, cz, cy], 1)),
                _dot(normal, np.stack([cz, zero, cx], 1)),
                _dot(normal, np.stack([one, zero, zero], 1)),
                _dot(normal, np.stack([zero, one, zero], 1)),
                _dot(normal, np.stack([zero, zero, one], 1)),
                zero,
            ]
            srow = np.stack(rows, 1).astype(F)
            phi[self_, i, :] = srow[self_]
    r = _normalized(refl, base, h, mn)
    rx, ry, rz = r[:, 0], r[:, 1], r[:, 2]
    phr = np.stack([rx * ry * rz, rx * ry, ry * rz, rx * rz, rx, ry, rz, one], 1).astype(F)
    return phi, base, normal, phr


def stencil_weights(phi, phr):
    """p = inv(Phi)^T phi(r) through the SVD, with f(s) = 1/s for s > 1e-6 and s otherwise
    (inverse_from_svd_kernel, svd.cu:129-148), in FP64; condition number max S / min S in FP32."""
    P = phi.astype(np.float64)
    U, S, Vt = np.linalg.svd(P)
    Sf = S.astype(F)
    cond = Sf.max(axis=1) / Sf.min(axis=1)
    fS = np.where(S > 1e-6, 1.0 / np.where(S > 0, S, 1.0), S)
    c = np.einsum("nij,nj->ni", Vt, phr.astype(np.float64))
    w = np.einsum("nij,nj->ni", U, fS * c)
    return w.astype(F), cond


def ghost_setup(cells, N, h, min_pos, threshold):
    """GhostCellSolver::precompute_ghost_matrix for one classification; returns a dict with
    the ghost cells' linear indices, order (1 = second order), weights, condition numbers, the
    9-slot rows (cols, vals) and the dense FP64 matrix."""
    typ = cells["type"].reshape(-1)
    gidx = cells["ghost_idx"].reshape(-1)
    ghosts = np.nonzero(typ == GHOST)[0]
    n = ghosts.shape[0]
    gcell = np.zeros(n, np.int64)
    gcell[gidx[ghosts]] = ghosts
    mn = np.asarray(min_pos, F)
    gc = np.stack([gcell // (N * N), (gcell // N) % N, gcell % N], 1).astype(np.int32)
    near = cells["point"].reshape(-1, 3)[gcell].astype(F)
    refl = cells["reflect"].reshape(-1, 3)[gcell].astype(F)
    phi, base, normal, phr = stencil_matrices(gc, near, refl, h, mn)
    if threshold > 0:
        w, cond = stencil_weights(phi, phr)
        second = ~(cond > F(threshold)) & ~np.any(np.isnan(w), axis=1)
    else:
        w = np.zeros((n, 8), F)
        cond = np.zeros(n, F)
        second = np.zeros(n, bool)
    w = np.where(second[:, None], w, F(0)).astype(F)
    cols = np.full((n, 9), -1, np.int32)
    vals = np.zeros((n, 9), F)
    cols[:, 8] = np.arange(n)
    vals[:, 8] = 1
    for i in range(8):
        nb = base + CORNERS[i]
        ok = np.all((nb >= 0) & (nb < N), axis=1) & second
        l = np.where(ok, (nb[:, 0].astype(np.int64) * N + nb[:, 1]) * N + nb[:, 2], 0)
        sel = ok & (l != gcell) & (typ[l] == GHOST) & (-w[:, i] != 0)
        cols[sel, i] = gidx[l[sel]]
        vals[sel, i] = -w[sel, i]
    A = np.zeros((n, n))
    for k in range(9):
        m = cols[:, k] >= 0
        A[np.nonzero(m)[0], cols[m, k]] += vals[m, k].astype(np.float64)
    inside = np.all((base >= 0) & (base + 1 < N), axis=1)  # the device reports these as out of grid
    return {"n": n, "cells": gcell, "order": second.astype(np.int32), "weights": w, "cond": cond, "inside": inside,
            "cols": cols, "vals": vals, "A": A, "base": base, "phi": phi, "phr": phr, "normal": normal,
            "threshold": threshold, "N": N, "h": h, "min": mn}


D6 = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.int32)


def ghost_rhs(setup, cells, p, neu):
    """construct_equation_kernel<false, true> (ghost_cell.cu:352-470): b per ghost from the
    pressure p (N^3 float32, the current level) and the Neumann data per triangle."""
    N, h = setup["N"], F(setup["h"])
    typ = cells["type"].reshape(-1)
    gidx = cells["ghost_idx"].reshape(-1)
    pf = p.reshape(-1).astype(F)
    gcell = setup["cells"]
    n = setup["n"]
    nv = neu.astype(F)[cells["nearest"].reshape(-1)[gcell]]
    gc = np.stack([gcell // (N * N), (gcell // N) % N, gcell % N], 1)
    b = np.zeros(n, F)
    first = setup["order"] == 0
    # first order: mean over the AIR face neighbours of (-a_n h + p)
    acc = np.zeros(n, F)
    cnt = np.zeros(n, np.int32)
    for d in D6:
        q = gc + d
        ok = np.all((q >= 0) & (q < N), axis=1)
        l = np.where(ok, (q[:, 0] * N + q[:, 1]) * N + q[:, 2], 0)
        air = ok & (typ[l] == AIR)
        acc = np.where(air, acc + ((-nv) * h + pf[l]), acc).astype(F)
        cnt += air
    with np.errstate(divide="ignore", invalid="ignore"):
        bf = (acc / cnt.astype(F)).astype(F)
    b[first] = bf[first]
    # second order
    w = setup["weights"]
    dist = cells["dist"].reshape(-1)[gcell].astype(F)
    bv = F(0) + ((-nv) * dist) * F(2)
    bv = bv.astype(F)
    scale = h / F(2)
    base = setup["base"]
    for i in range(8):
        nb = base + CORNERS[i]
        ok = np.all((nb >= 0) & (nb < N), axis=1)
        l = np.where(ok, (nb[:, 0].astype(np.int64) * N + nb[:, 1]) * N + nb[:, 2], 0)
        is_self = ok & (gidx[l] == np.arange(n)) & (typ[l] == GHOST)
        other = ok & ~is_self & (typ[l] != GHOST)
        bv = np.where(is_self, bv + w[:, i] * (scale * nv), bv).astype(F)
        bv = np.where(other, bv + w[:, i] * pf[l], bv).astype(F)
    b[~first] = bv[~first]
    return b


def ghost_solve(setup, b):
    """The ghost values: b itself when the threshold is 0, else the solution of A x = b (FP64,
    zero when b is zero: BiCGSTAB_Solver::solve, sparse.cu:512-545)."""
    if setup["threshold"] <= 0:
        return b.astype(F)
    if not np.any(b):
        return np.zeros_like(b, dtype=F)
    return np.linalg.solve(setup["A"], b.astype(np.float64)).astype(F)


def ghost_step(oracle, setup, cells, neu):
    """GhostCellSolver::update (ghost_cell.h:131-141) with the C++ oracle's FDTD: step, right-hand
    side at the new level, solve, ghost values in, SOLID cells to zero. Returns (b, x)."""
    oracle.fdtd_step()
    p = oracle.grid(0).reshape(-1).copy()
    b = ghost_rhs(setup, cells, p, neu)
    x = ghost_solve(setup, b)
    p[setup["cells"]] = x
    p[cells["type"].reshape(-1) == SOLID] = 0
    oracle.set_grid(0, p)
    return b, x

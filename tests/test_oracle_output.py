"""Pins of the oracle's FFAT map collection (collect_ffat_map, experiments/static/main.cu:60-80)
against a slice statement of the same six faces, at the widths the reference's experiments use."""
import numpy as np
import pytest

import oracle as oracle_mod


def _faces_by_slices(g, w):
    """The six faces of the cube of half width w around c = N/2, as array slices: face b of axis
    i and sign s holds p at c + s w e_i + (x - w) e_{i+1} + (y - w) e_{i+2} (axes cyclic)."""
    c = g.shape[0] // 2
    lo, hi = c - w, c + w
    out = []
    # axis 0: (x, y) -> (y axis, z axis)
    for s in (-1, 1):
        out.append(g[c + s * w, lo:hi, lo:hi])
    # axis 1: (x, y) -> (z axis, x axis): p[x_ax = lo + y, y_ax = c + s w, z_ax = lo + x]
    for s in (-1, 1):
        out.append(g[lo:hi, c + s * w, lo:hi].T)
    # axis 2: (x, y) -> (x axis, y axis)
    for s in (-1, 1):
        out.append(g[lo:hi, lo:hi, c + s * w])
    return np.abs(np.stack(out))


@pytest.mark.parametrize("N,w", [(32, 9), (32, 6), (32, 11), (33, 10), (64, 19), (24, 11)])
def test_ffat_collect_matches_face_slices(N, w):
    rng = np.random.default_rng(N * 100 + w)
    g = rng.standard_normal((N, N, N)).astype(np.float32)
    m = np.zeros((6, 2 * w, 2 * w), np.float32)
    oracle_mod.ffat_collect(g, w, m)
    assert np.array_equal(m, _faces_by_slices(g, w))


def test_ffat_collect_accumulates():
    """Two collections of the same grid: every entry is 2|p| (exact in binary floating point)."""
    rng = np.random.default_rng(5)
    g = rng.standard_normal((32, 32, 32)).astype(np.float32)
    m = np.zeros((6, 18, 18), np.float32)
    oracle_mod.ffat_collect(g, 9, m)
    oracle_mod.ffat_collect(g, 9, m)
    assert np.array_equal(m, 2 * _faces_by_slices(g, 9))


def test_ffat_faces_are_the_cube_boundary():
    """A grid holding 1 on the boundary of the cube [c-w, c+w]^3 and 0 elsewhere: every map entry
    is 1 (each face cell lies on the boundary), so no face reads off the cube."""
    N, w = 32, 7
    c = N // 2
    g = np.zeros((N, N, N), np.float32)
    cube = np.zeros_like(g, dtype=bool)
    cube[c - w:c + w + 1, c - w:c + w + 1, c - w:c + w + 1] = True
    inner = np.zeros_like(cube)
    inner[c - w + 1:c + w, c - w + 1:c + w, c - w + 1:c + w] = True
    g[cube & ~inner] = 1.0
    m = np.zeros((6, 2 * w, 2 * w), np.float32)
    oracle_mod.ffat_collect(g, w, m)
    assert np.all(m == 1.0)


def test_reference_experiment_widths():
    """The widths the experiments use fit the grid: static, w = int(size * res) for sizes
    0.20..0.35; acoustic, w = int(0.3 N)."""
    for res in (32, 48, 64, 128):
        for size in (0.20, 0.25, 0.30, 0.35):
            w = int(np.float32(size) * res)
            assert 1 <= w and res // 2 + w <= res - 1
        assert res // 2 + int(res * 0.3) <= res - 1

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


# ------------------------------------------------------------------ slice images
# (save_grid, src/utils/visualize.h:120-132; JetColor and the image passes, src/ui/window.cu)
N_IMG = 40                      # upsampling 1080 // 40 = 27, W = 1080
F_IMG = 1080 // N_IMG
MN_IMG = (-0.1, -0.1, -0.1)
H_IMG = 0.005


def _interior_pixels(W, f):
    """PNG pixels at least 2 px from every cell border (no grid line) inside the interior cells."""
    x = np.arange(W)
    far = np.abs((x + 0.5) - ((x // f) * f + f / 2.0)) < f / 2.0 - 2.0
    inner = (x // f > 0) & (x // f < N_IMG - 1)
    return far & inner


@pytest.mark.parametrize("v,expect", [(-1.0, (0, 0, 127)), (-0.5, (0, 127, 127)), (0.0, (0, 127, 0)),
                                      (0.5, (127, 127, 0)), (1.0, (127, 0, 0)), (3.0, (127, 0, 0))])
def test_slice_jet_anchors(v, expect):
    """A uniform plane shows the jet colour of its value at half brightness: blue, cyan,
    green, yellow, red at -1, -1/2, 0, 1/2, 1 of the range (clamped beyond)."""
    g = np.full((N_IMG,) * 3, v, np.float32)
    img = oracle_mod.slice_image(g, H_IMG, MN_IMG, 1, N_IMG // 2, 1.0)
    m = _interior_pixels(img.shape[0], F_IMG)
    px = img[np.ix_(m[::-1], m)]
    assert np.all(px[..., :3] == np.array(expect, np.uint8)) and np.all(px[..., 3] == 255)


def test_slice_cell_borders_are_blended_white():
    """Pixels within 1.5 px of a cell border mix 40 % white into the colour: green 127 becomes
    0.6 * 127 + 102 = 178 (truncated), the zero channels 102."""
    g = np.zeros((N_IMG,) * 3, np.float32)
    img = oracle_mod.slice_image(g, H_IMG, MN_IMG, 2, 7, 1.0)
    W = img.shape[0]
    x = np.arange(W)
    border = (F_IMG / 2.0 - np.abs((x + 0.5) - ((x // F_IMG) * F_IMG + F_IMG / 2.0))) <= 1.5
    row = img[W // 2 + 3, :, :3]  # a PNG row away from horizontal borders
    assert np.all(row[border] == np.array([102, 178, 102], np.uint8))
    assert np.all(row[~border] == np.array([0, 127, 0], np.uint8))


def test_slice_orientation():
    """Plane axis 2: the image's first axis is x (PNG columns, left to right) and its second
    axis y (PNG rows, bottom to top): a field red for x >= N/2 and y >= N/2 fills the top-right
    quadrant."""
    g = np.full((N_IMG,) * 3, -1.0, np.float32)
    g[N_IMG // 2:, N_IMG // 2:, :] = 1.0
    img = oracle_mod.slice_image(g, H_IMG, MN_IMG, 2, 5, 1.0)
    W = img.shape[0]
    lo, hi = W // 4 + F_IMG // 2, 3 * W // 4 + F_IMG // 2  # cell centres, away from borders
    assert tuple(img[lo, hi, :3]) == (127, 0, 0)  # top right
    assert tuple(img[hi, lo, :3]) == (0, 0, 127)  # bottom left
    assert tuple(img[lo, lo, :3]) == (0, 0, 127)  # top left
    assert tuple(img[hi, hi, :3]) == (0, 0, 127)  # bottom right


def test_slice_linear_field_is_interpolated_monotonically():
    """A field rising linearly in x within the blue-to-cyan band: the green channel grows along
    a PNG row (bilinear interpolation between cell centres), from about 0 to about 127."""
    x = np.arange(N_IMG, dtype=np.float32)
    g = np.broadcast_to((-1.0 + 0.45 * x / (N_IMG - 1))[:, None, None], (N_IMG,) * 3).astype(np.float32).copy()
    img = oracle_mod.slice_image(g, H_IMG, MN_IMG, 2, 9, 1.0)
    m = _interior_pixels(img.shape[0], F_IMG)
    green = img[img.shape[0] // 2 + 5, m, 1].astype(int)
    assert np.all(np.diff(green) >= 0) and green[0] < 10 and green[-1] > 100


def test_slice_mesh_cut_is_drawn_white():
    """One triangle crossing plane z (axis 2): pixels on its cut are pure white, pixels 20 px
    away from it keep the field colour."""
    g = np.zeros((N_IMG,) * 3, np.float32)
    idx = 20
    zc = MN_IMG[2] + (idx - N_IMG / 2.0 + 0.5) * H_IMG + 0.1  # plane through the grid centre offset
    verts = np.array([[-0.05, -0.04, zc - 0.03], [0.06, -0.04, zc - 0.03], [0.0, 0.05, zc + 0.03]], np.float32)
    tris = np.array([[0, 1, 2]], np.int32)
    img = oracle_mod.slice_image(g, H_IMG, MN_IMG, 2, idx, 1.0, verts, tris)
    W = img.shape[0]
    # the cut at the plane's height: the edges 0-2 and 1-2 at half height, y = 0.005
    zp = np.float32(MN_IMG[2] + N_IMG * H_IMG / 2.0) + np.float32(idx - N_IMG / 2.0 + 0.5) * np.float32(H_IMG)
    t = (zp - verts[0, 2]) / (verts[2, 2] - verts[0, 2])
    a = verts[0] + t * (verts[2] - verts[0])
    b = verts[1] + t * (verts[2] - verts[1])
    for s in (0.2, 0.5, 0.8):
        p = a + s * (b - a)
        u = (p[0] - MN_IMG[0]) / (N_IMG * H_IMG) * W
        v = (p[1] - MN_IMG[1]) / (N_IMG * H_IMG) * W
        assert tuple(img[W - 1 - int(v), int(u)]) == (255, 255, 255, 255)
        assert tuple(img[W - 1 - int(v) - 20, int(u)]) != (255, 255, 255, 255)

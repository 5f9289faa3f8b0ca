"""GPU FFAT map collection (pppm_ffat_collect through the C ABI) against the oracle's
collect_ffat_map on the same grids: the static experiment's four widths and the acoustic
experiment's, accumulated over the steps after a warm-up, bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_ffat_maps_bitexact(amd, oracle_mod, scenes_mod):
    import torch
    sc = scenes_mod.c1_static_sphere(steps=60)
    N = sc.res
    widths = sorted({int(np.float32(f) * N) for f in (0.20, 0.25, 0.30, 0.35)} | {int(N * 0.3)})
    print(f"ffat: icosphere L3 on {N}^3, widths {widths}", flush=True)
    g = amd.PPPMSolver(sc.res, sc.dl, sc.dt, sc.min_pos, 0, 0.83, sc.c)
    g.set_mesh(sc.verts, sc.tris)
    maps = {w: torch.zeros((6, 2 * w, 2 * w), dtype=torch.float32, device="cuda") for w in widths}
    ref = {w: np.zeros((6, 2 * w, 2 * w), np.float32) for w in widths}
    torch.cuda.synchronize()  # the maps are cleared on torch's stream, collected on the solver's
    for i in range(sc.steps):
        g.update_grid_and_face(sc.neumann[i])
        if i > 20:
            for w in widths:
                g.ffat_collect(w, maps[w])
            torch.cuda.synchronize()
            grid = g.get_grid(0).reshape(N, N, N)
            for w in widths:
                oracle_mod.ffat_collect(grid, w, ref[w])
    torch.cuda.synchronize()
    for w in widths:
        got = maps[w].cpu().numpy()
        assert np.abs(got).max() > 0
        assert np.array_equal(got.view(np.uint32), ref[w].view(np.uint32)), w


def test_ffat_width_outside_grid_is_rejected(amd):
    import torch
    g = amd.PPPMSolver(32, 0.005, 1e-5, (0, 0, 0), 0, 0.83, 343.0)
    m = torch.zeros((6, 32, 32), dtype=torch.float32, device="cuda")
    with pytest.raises(amd.PPPMError):
        g.ffat_collect(16, m)


@pytest.mark.parametrize("axis,index", [(0, 16), (1, 16), (2, 11), (0, 1)])
def test_slice_image_bitexact(amd, oracle_mod, scenes_mod, axis, index):
    """save_grid's picture (colour map, interpolation, cell borders, the mesh's cut) of a plane
    of the C1 scene after 30 steps, equal to the oracle's byte for byte; written and read back
    as a PNG file."""
    import os
    import tempfile
    sc = scenes_mod.c1_static_sphere(steps=30)
    print(f"slice image: axis {axis}, index {index}", flush=True)
    g = amd.PPPMSolver(sc.res, sc.dl, sc.dt, sc.min_pos, 0, 0.83, sc.c)
    g.set_mesh(sc.verts, sc.tris)
    for i in range(sc.steps):
        g.update_grid_and_face(sc.neumann[i])
    vmax = float(np.abs(g.get_grid(0)).max()) * 0.5
    img = g.slice_image(axis, index, vmax)
    ref = oracle_mod.slice_image(g.get_grid(0), sc.dl, sc.min_pos, axis, index, vmax, sc.verts, sc.tris)
    diff = int(np.count_nonzero(np.any(img != ref, axis=2)))
    print(f"pixels differing: {diff} of {img.shape[0] * img.shape[1]}", flush=True)
    assert diff == 0
    path = os.path.join(tempfile.mkdtemp(), "slice.png")
    amd.write_png(path, img)
    assert np.array_equal(amd.read_png(path), img)

"""pppm_update_mesh_and_step enqueues the step behind the mesh update and reads the update's
status while the step runs; the step's kernels do nothing when the update raised an error bit.
Its recovery paths against the separate calls (update_mesh, then update_grid_and_face) on a
second handle, bit for bit:
- an update whose face lists outgrow their capacity (the sphere shrunk to a quarter: the 4x4x4
  blocks hold many more triangles) and whose face-cache pool overflows: grown, updated again,
  stepped again;
- an update with every triangle outside the grid: the error is returned, the time index and the
  grids stay as they were, and the next valid update steps as if the failed one never happened."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(amd, sc):
    hs = []
    for _ in range(2):
        g = amd.PPPMSolver(sc.res, sc.dl, sc.dt, sc.min_pos, 0, 0.83, sc.c)
        g.set_mesh(sc.verts, sc.tris)
        hs.append(g)
    return hs


def _same_grids(a, b):
    for rel in (0, -1):
        assert np.array_equal(a.get_grid(rel).view(np.uint32), b.get_grid(rel).view(np.uint32)), rel


def test_overflowing_update_is_grown_and_stepped_again(amd, scenes_mod):
    sc = scenes_mod.c1_static_sphere(steps=12)
    print("update_mesh_and_step overflow: C1 sphere shrunk to 1/4, 12 steps", flush=True)
    a, b = _pair(amd, sc)
    ctr = sc.verts.mean(axis=0)
    small = (ctr + (sc.verts - ctr) * np.float32(0.25)).astype(np.float32)
    p4_before = a.stats()["P4"]
    for i in range(12):
        v = sc.verts if i < 3 or i >= 8 else small
        a.update_mesh_and_step(v, sc.neumann[i])
        b.update_mesh(v)
        b.update_grid_and_face(sc.neumann[i])
        assert a.time_idx() == b.time_idx() == i
        _same_grids(a, b)
        if i == 3:
            assert a.stats()["P4"] > 4 * p4_before  # the shrunk sphere overflowed the initial capacity
    assert np.abs(a.get_grid(0)).max() > 0


def test_invalid_update_skips_the_step(amd, scenes_mod):
    sc = scenes_mod.c1_static_sphere(steps=8)
    print("update_mesh_and_step with an invalid mesh: C1 sphere, 8 steps", flush=True)
    a, b = _pair(amd, sc)
    for i in range(3):
        a.update_mesh_and_step(sc.verts, sc.neumann[i])
        b.update_mesh_and_step(sc.verts, sc.neumann[i])
    g0, g1 = a.get_grid(0), a.get_grid(-1)
    bad = sc.verts.copy()
    bad[:, 0] += np.float32(1.0)
    with pytest.raises(amd.PPPMError):
        a.update_mesh_and_step(bad, sc.neumann[3])
    assert a.time_idx() == 2
    assert np.array_equal(a.get_grid(0).view(np.uint32), g0.view(np.uint32))
    assert np.array_equal(a.get_grid(-1).view(np.uint32), g1.view(np.uint32))
    with pytest.raises(amd.PPPMError, match="last mesh update failed"):
        a.update_grid_and_face(sc.neumann[3])
    for i in range(3, 8):
        a.update_mesh_and_step(sc.verts, sc.neumann[i])
        b.update_mesh_and_step(sc.verts, sc.neumann[i])
        _same_grids(a, b)

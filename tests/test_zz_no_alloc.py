"""The allocation-free entry point pppm_run_noalloc against pppm_run (eager), through the C ABI:
two calls in a row on two handles, the second with a smaller mesh and grid, so the second call
runs on the scratch array the first one left behind. Listener samples and both grid levels
must agree bit for bit."""
import numpy as np
import pytest


def _reference(amd, sc, steps, listeners):
    import torch
    g = amd.PPPMSolver(sc.res, sc.dl, sc.dt, sc.min_pos, 0, 0.83, sc.c)
    g.set_mesh(sc.verts, sc.tris)
    out = torch.zeros(steps, len(listeners), device="cuda")
    g.run(torch.from_numpy(sc.neumann[:steps]).cuda(), steps, listeners, out, use_graph=False)
    torch.cuda.synchronize()
    return out.cpu().numpy(), [g.get_grid(rel) for rel in (0, -1)]


@pytest.mark.gpu
def test_run_noalloc_equals_run(amd, scenes_mod):
    import torch
    big = scenes_mod.c1_static_sphere(steps=150)
    small = scenes_mod.c3_scene(24, 1, seed=24, steps=90)
    cases = [(big, 150, [big.listener, (16, 16, 9), (3, 4, 5)]),
             (small, 90, [small.listener, (12, 12, 12)])]
    handles = []
    for sc, steps, lis in cases:
        print(f"run_noalloc vs run: {sc.name} {sc.res}^3 T={len(sc.tris)} steps={steps}", flush=True)
        g = amd.PPPMSolver(sc.res, sc.dl, sc.dt, sc.min_pos, 0, 0.83, sc.c)
        g.set_mesh(sc.verts, sc.tris)
        out = torch.full((steps, len(lis)), float("nan"), device="cuda")
        g.run_noalloc(torch.from_numpy(sc.neumann[:steps]).cuda(), steps, lis, out)
        torch.cuda.synchronize()
        ref_out, ref_grids = _reference(amd, sc, steps, lis)
        got = out.cpu().numpy()
        assert np.array_equal(got.view(np.uint32), ref_out.view(np.uint32))
        for rel, rg in zip((0, -1), ref_grids):
            assert np.array_equal(g.get_grid(rel).view(np.uint32), rg.view(np.uint32)), rel
        assert np.abs(got).max() > 0
        handles.append(g)


@pytest.mark.gpu
def test_run_noalloc_rejects_a_mesh_beyond_the_scratch(amd, scenes_mod):
    """Capacity check before any launch: 120 000 small triangles need more than 16 MiB."""
    import torch
    rng = np.random.default_rng(5)
    n, N, h = 120000, 96, 0.005
    off = (rng.random((n, 3)) * (N - 8) + 4) * h
    verts = np.repeat(off, 3, axis=0).astype(np.float32)
    verts[1::3, 0] += 0.1 * h
    verts[2::3, 1] += 0.1 * h
    tris = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
    g = amd.PPPMSolver(N, h, h / (np.sqrt(3) * 343.0), (0, 0, 0), 0, 0.83, 343.0)
    g.set_mesh(verts, tris)
    with pytest.raises(amd.PPPMError, match="scratch"):
        g.run_noalloc(torch.zeros(n, device="cuda"), 1, [], None)

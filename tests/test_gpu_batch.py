"""Batched scenes (pppm_batch_*) against the scenes' own single-scene runs, through the C ABI:
every scene's listener samples and both live grid levels bit for bit, for every FDTD interior
variant (one thread per cell and the four LDS-staged tile shapes), with and without the graph,
and after the autotuner has run on the same scenes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (grid edge, icosphere level): uneven sizes, including edges that are not multiples of a tile
CASES = [(24, 1), (37, 2), (48, 2), (61, 3), (80, 3), (96, 4)]


def _scenes(scenes_mod, steps, cases=CASES):
    scs = [scenes_mod.c3_scene(n, lv, seed=n, steps=steps) for n, lv in cases]
    # listener two cells past the object's largest half-extent along x (the scenes' own corner
    # listener is more than 48 stencil steps from the surface on the larger grids)
    for sc in scs:
        sc.listener = (3 * sc.res // 4 + 2, sc.res // 2, sc.res // 2)
    return scs


def _single(amd, sc, steps):
    import torch
    g = amd.PPPMSolver(sc.res, sc.dl, sc.dt, sc.min_pos, 0, 0.83, sc.c)
    g.set_mesh(sc.verts, sc.tris)
    out = torch.zeros(steps, 1, device="cuda")
    g.run(torch.from_numpy(sc.neumann[:steps]).cuda(), steps, [sc.listener], out, use_graph=True)
    torch.cuda.synchronize()
    return out[:, 0].cpu().numpy(), g.get_grid(0), g.get_grid(-1)


def _batch(amd, scs, steps, variants=None, use_graph=True, autotune=False):
    import torch
    hs = []
    for sc in scs:
        g = amd.PPPMSolver(sc.res, sc.dl, sc.dt, sc.min_pos, 0, 0.83, sc.c)
        g.set_mesh(sc.verts, sc.tris)
        hs.append(g)
    b = amd.PPPMBatch(hs)
    if autotune:
        table = b.autotune(reps=3)
        assert table and all(len(v) == 5 for v in table.values())
    if variants is not None:
        b.set_variants(variants)
    ns = [torch.from_numpy(sc.neumann[:steps]).cuda() for sc in scs]
    out = torch.zeros(steps, len(scs), device="cuda")
    b.run(ns, steps, [sc.listener for sc in scs], out, use_graph=use_graph)
    torch.cuda.synchronize()
    return out.cpu().numpy(), hs, b


def _same(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4])
def test_batch_equals_single_runs(amd, scenes_mod, variant):
    steps = 48
    scs = _scenes(scenes_mod, steps)
    print(f"batch vs single: {len(scs)} scenes, FDTD variant {variant}, {steps} steps", flush=True)
    out, hs, _ = _batch(amd, scs, steps, variants=[variant] * len(scs))
    for i, sc in enumerate(scs):
        ref, g0, g1 = _single(amd, sc, steps)
        assert _same(out[:, i], ref), (i, sc.name)
        assert _same(hs[i].get_grid(0), g0) and _same(hs[i].get_grid(-1), g1), (i, sc.name)
        assert np.abs(ref).max() > 0 and np.abs(g0).max() > 0, (i, sc.name)


def test_batch_mixed_variants_eager_and_autotuned(amd, scenes_mod):
    """Mixed variants in one batch, eager launches; then a batch that ran the autotuner first
    (its timing launches must leave the scenes' state as it was); both equal the single runs.
    The second run continues from the first (t advanced), as consecutive pppm_run calls do."""
    steps = 40
    scs = _scenes(scenes_mod, 2 * steps)
    print(f"batch mixed variants eager + autotuned: {len(scs)} scenes, {steps} + {steps} steps", flush=True)
    out_e, hs_e, b = _batch(amd, scs, steps, variants=[0, 1, 2, 3, 4, 1], use_graph=False)
    out_t, hs_t, _ = _batch(amd, scs, steps, autotune=True)
    for i, sc in enumerate(scs):
        ref, g0, g1 = _single(amd, sc, steps)
        assert _same(out_e[:, i], ref) and _same(out_t[:, i], ref), (i, sc.name)
        assert _same(hs_t[i].get_grid(0), g0) and _same(hs_t[i].get_grid(-1), g1)
    # continue the eager batch: steps 40..79 equal a single 80-step run's second half
    import torch
    ns = [torch.from_numpy(np.ascontiguousarray(sc.neumann[steps:2 * steps])).cuda() for sc in scs]
    out2 = torch.zeros(steps, len(scs), device="cuda")
    b.run(ns, steps, [sc.listener for sc in scs], out2, use_graph=True)
    out2 = out2.cpu().numpy()
    for i, sc in enumerate(scs):
        ref, _, _ = _single(amd, sc, 2 * steps)
        assert _same(out2[:, i], ref[steps:]), (i, sc.name)


def test_batch_several_listeners_per_scene(amd, scenes_mod):
    """Three listeners per scene (one near the object, one in the corrected region's corner case
    of the scene's own corner, one on the axis): each scene's samples equal its single run's."""
    import torch
    steps = 40
    scs = _scenes(scenes_mod, steps, CASES[:4])
    lis = [[sc.listener, (sc.res // 2, sc.res // 2, 3 * sc.res // 4 + 1), (2, 3, 4)] for sc in scs]
    print(f"batch, 3 listeners per scene: {len(scs)} scenes, {steps} steps", flush=True)
    hs = []
    for sc in scs:
        g = amd.PPPMSolver(sc.res, sc.dl, sc.dt, sc.min_pos, 0, 0.83, sc.c)
        g.set_mesh(sc.verts, sc.tris)
        hs.append(g)
    b = amd.PPPMBatch(hs)
    out = torch.zeros(steps, len(scs), 3, device="cuda")
    b.run([torch.from_numpy(sc.neumann[:steps]).cuda() for sc in scs], steps, lis, out)
    out = out.cpu().numpy()
    for i, sc in enumerate(scs):
        g = amd.PPPMSolver(sc.res, sc.dl, sc.dt, sc.min_pos, 0, 0.83, sc.c)
        g.set_mesh(sc.verts, sc.tris)
        ref = torch.zeros(steps, 3, device="cuda")
        g.run(torch.from_numpy(sc.neumann[:steps]).cuda(), steps, lis[i], ref)
        assert _same(out[:, i, :], ref.cpu().numpy()), (i, sc.name)
    assert np.abs(out[:, :, :2]).max() > 0

"""The GPU adjoint (pppm_adjoint_*, through the C ABI and the torch.autograd.Function) against the
oracle's FP64 adjoint, which central finite differences pin (tests/test_oracle_adjoint.py):
- small scenes: the gradient of random listener cotangents, every entry;
- the adjoint configuration (48^3, icosphere L3, 2000 steps, scenes.c4_adjoint) against the
  committed FP64 reference (tests/golden/make_c4_adjoint.py): the loss, 4096 sampled gradient
  entries, and central differences (eps = 1e-3, FP64) at 32 random (step, triangle) entries;
- the checkpointed loss (sqrt(T) segments, recomputation) gives the same gradient as the
  one-sweep signal gradient, bit for bit.
Tolerances are relative to the largest gradient entry; FP32 stepping over 2000 steps."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _solver(amd, sc):
    g = amd.PPPMSolver(sc.res, sc.dl, sc.dt, sc.min_pos, 0, 0.83, sc.c)
    g.set_mesh(sc.verts, sc.tris)
    return g


@pytest.mark.parametrize("n,lv", [(24, 1), (20, 2), (32, 3)])
def test_adjoint_matches_the_fp64_adjoint(amd, oracle_mod, scenes_mod, n, lv):
    import torch
    T = 160
    sc = scenes_mod.c3_scene(n, lv, seed=n, steps=T)
    print(f"adjoint vs FP64 adjoint: {n}^3 L{lv}, {T} steps", flush=True)
    o = oracle_mod.Oracle(sc.res, sc.dl, sc.dt, sc.min_pos, sc.c, 0.83, oracle_mod.FP64)
    o.set_mesh(sc.verts, sc.tris)
    tri = o.triangles()
    lis = [sc.listener, tuple(int(x) for x in tri[0, 7:10] + np.array([1, 0, 0]))]
    rng = np.random.default_rng(n)
    gout = rng.standard_normal((T, len(lis)))
    ref = o.linear_adjoint(gout, lis)
    g = _solver(amd, sc)
    grad = torch.full((T, len(sc.tris)), float("nan"), device="cuda")
    g.adjoint_reset(T - 1)
    # two segments: the adjoint state carries across calls
    g.adjoint_run(torch.from_numpy(gout[T // 2:].astype(np.float32)).cuda(), T - T // 2, lis, grad[T // 2:])
    g.adjoint_run(torch.from_numpy(gout[:T // 2].astype(np.float32)).cuda(), T // 2, lis, grad[:T // 2],
                  use_graph=False)
    got = grad.cpu().numpy().astype(np.float64)
    scale = np.max(np.abs(ref))
    err = np.max(np.abs(got - ref)) / scale
    print(f"  max |grad - FP64| / max |FP64| = {err:.2e}")
    assert scale > 0 and np.all(np.isfinite(got)) and err < 1e-4


def test_c4_gradient_against_fd_and_fp64(amd, scenes_mod):
    import torch
    import autograd as pad
    ref = np.load(os.path.join(HERE, "golden", "c4_adjoint.npz"))
    sc = scenes_mod.c4_adjoint()
    T = sc.steps
    print(f"C4 adjoint: {sc.res}^3, {len(sc.tris)} triangles, {T} steps", flush=True)
    s = _solver(amd, sc)
    lis = [sc.listener]
    target = torch.from_numpy(ref["target"].astype(np.float32)).cuda()
    g = torch.from_numpy(sc.neumann).cuda().requires_grad_(True)
    J = pad.listener_loss(g, target, s, lis)
    J.backward()
    gc = g.grad.detach().cpu().numpy().astype(np.float64)
    scale = float(ref["grad_maxabs"])
    loss_err = abs(float(J) - float(ref["loss"])) / float(ref["loss"])
    samp = np.max(np.abs(gc[ref["sample_steps"], ref["sample_tris"]] - ref["sample_grad"])) / scale
    fd = np.max(np.abs(gc[ref["fd_steps"], ref["fd_tris"]] - ref["fd"])) / scale
    fd_adj = np.max(np.abs(ref["fd_adjoint"] - ref["fd"])) / scale
    print(f"  loss rel err {loss_err:.2e}; grad vs FP64 adjoint (4096 entries) {samp:.2e}; grad vs FD (32) {fd:.2e};"
          f" FP64 adjoint vs FD {fd_adj:.2e}")
    assert fd_adj < 2e-5  # FP64 rounding of J over 2000 steps: 1.9e-6 measured
    assert loss_err < 1e-4 and samp < 1e-3 and fd < 1e-3
    # the one-sweep gradient of the signal equals the checkpointed one bit for bit
    g2 = torch.from_numpy(sc.neumann).cuda().requires_grad_(True)
    out = pad.listener_signal(g2, s, lis)
    ((out - target) ** 2).sum().backward()
    assert torch.equal(g2.grad, g.grad)


def test_state_checkpoint_round_trip(amd, scenes_mod):
    """save_state / load_state: a run resumed from a checkpoint equals the uninterrupted run."""
    import torch
    sc = scenes_mod.c3_scene(37, 2, seed=37, steps=60)
    a, b = _solver(amd, sc), _solver(amd, sc)
    nm = torch.from_numpy(sc.neumann).cuda()
    oa = torch.zeros(60, 1, device="cuda")
    a.run(nm, 60, [sc.listener], oa)
    ob = torch.zeros(60, 1, device="cuda")
    b.run(nm[:25], 25, [sc.listener], ob[:25])
    ck = torch.empty(b.state_bytes(), dtype=torch.uint8, device="cuda")
    b.save_state(ck)
    b.run(nm[25:40], 15, [sc.listener], ob[25:40])   # diverge, then restore
    b.load_state(ck)
    assert b.time_idx() == 24
    b.run(nm[25:], 35, [sc.listener], ob[25:])
    assert torch.equal(oa, ob)
    for rel in (0, -1):
        assert np.array_equal(a.get_grid(rel), b.get_grid(rel))


def test_segment_runs_replay_one_graph(amd, scenes_mod):
    """Runs over consecutive segments of one series replay the graph captured for the first one
    (their launch arguments differ only by the rows of the series and of the output): samples and
    grid equal an eager run, and a second pass into another output buffer equals the first."""
    import torch
    T, S = 120, 20
    sc = scenes_mod.c3_scene(37, 2, seed=11, steps=T)
    a, b = _solver(amd, sc), _solver(amd, sc)
    nm = torch.from_numpy(sc.neumann).cuda()
    oa = torch.zeros(T, 1, device="cuda")
    a.run(nm, T, [sc.listener], oa, use_graph=False)
    for rep in range(2):
        print(f"segment pass {rep}", flush=True)
        ob = torch.zeros(T, 1, device="cuda")
        b.reset()
        for k in range(T // S):
            b.run(nm[S * k:S * (k + 1)], S, [sc.listener], ob[S * k:S * (k + 1)], use_graph=True)
        assert torch.equal(oa, ob)
        for rel in (0, -1):
            assert np.array_equal(a.get_grid(rel), b.get_grid(rel))


def test_segment_runs_past_the_graph_cache(amd, scenes_mod):
    """Segments of six different lengths (more graphs than a handle keeps, so the oldest are
    evicted while the run goes on), run twice: both passes equal an eager run."""
    import torch
    lens = [5, 7, 9, 11, 13, 15]
    T = sum(lens)
    sc = scenes_mod.c3_scene(37, 2, seed=12, steps=T)
    a, b = _solver(amd, sc), _solver(amd, sc)
    nm = torch.from_numpy(sc.neumann).cuda()
    oa = torch.zeros(T, 1, device="cuda")
    a.run(nm, T, [sc.listener], oa, use_graph=False)
    for rep in range(2):
        print(f"segment lengths {lens}, pass {rep}", flush=True)
        ob = torch.zeros(T, 1, device="cuda")
        b.reset()
        t = 0
        for n in lens:
            b.run(nm[t:t + n], n, [sc.listener], ob[t:t + n], use_graph=True)
            t += n
        assert torch.equal(oa, ob)
        for rel in (0, -1):
            assert np.array_equal(a.get_grid(rel), b.get_grid(rel))

"""The batched adjoint (pppm_batch_adjoint_run) against each scene's own adjoint run, bit for
bit (uneven scenes, several listeners, graph and eager, reversed in two segments); the batched,
sqrt(T)-checkpointed listener loss (autograd.BatchListenerLoss: batched forward, recompute and
adjoint) against the single-scene checkpointed loss of each scene (autograd.ListenerLoss):
losses and gradients bit for bit; and the adjoint identity <y, F g> = <F^T y, g> on a scene of
the batched-adjoint configuration's kind (radially perturbed icosphere, 8 listeners)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_adjoint_equals_single_scene_adjoints(amd, scenes_mod, use_graph):
    import torch
    T, L = 70, 3
    cases = [(24, 1), (41, 3), (30, 2), (52, 2)]
    scs = [scenes_mod.c3_scene(n, lv, seed=n + 7, steps=T) for n, lv in cases]
    lis = [[(3 * sc.res // 4 + 2, sc.res // 2, sc.res // 2), (sc.res // 2, 2 + sc.res // 8, sc.res // 2),
            (sc.res // 3, sc.res // 2 + 3, 2 + sc.res // 5)] for sc in scs]
    print(f"batched adjoint: {len(scs)} scenes, {T} steps, {L} listeners, graph {use_graph}", flush=True)

    def solver(sc):
        g = amd.PPPMSolver(sc.res, sc.dl, sc.dt, sc.min_pos, 0, 0.83, sc.c)
        g.set_mesh(sc.verts, sc.tris)
        return g

    sv = [solver(sc) for sc in scs]
    b = amd.PPPMBatch(sv)
    rng = np.random.default_rng(5)
    y = torch.from_numpy(rng.standard_normal((T, len(scs), L)).astype(np.float32)).cuda()
    grads = [torch.full((T, sc.tris.shape[0]), float("nan"), device="cuda") for sc in scs]
    b.adjoint_reset(T - 1)
    cut = 29  # two consecutive calls: steps [cut, T), then [0, cut)
    b.adjoint_run(y[cut:].contiguous(), T - cut, lis, [g[cut:] for g in grads], use_graph=use_graph)
    b.adjoint_run(y[:cut].contiguous(), cut, lis, [g[:cut] for g in grads], use_graph=use_graph)
    for i, sc in enumerate(scs):
        s1 = solver(sc)
        g1 = torch.zeros(T, sc.tris.shape[0], device="cuda")
        s1.adjoint_reset(T - 1)
        s1.adjoint_run(y[:, i].contiguous(), T, lis[i], g1, use_graph=use_graph)
        assert torch.equal(g1, grads[i]), i
        assert float(g1.abs().max()) > 0, i


def test_batch_loss_equals_single_scene_losses(amd, scenes_mod):
    import torch
    import autograd as pad
    T = 90
    cases = [(24, 1), (37, 2), (30, 2)]
    scs = [scenes_mod.c3_scene(n, lv, seed=n, steps=T) for n, lv in cases]
    lis = [[(3 * sc.res // 4 + 2, sc.res // 2, sc.res // 2), (sc.res // 2, 2 + sc.res // 8, sc.res // 2)]
           for sc in scs]
    print(f"batched checkpointed loss: {len(scs)} scenes, {T} steps, 2 listeners", flush=True)
    rng = np.random.default_rng(3)
    targets = [torch.from_numpy(rng.standard_normal((T, 2)).astype(np.float32) * 1e-4).cuda() for _ in scs]

    def solver(sc):
        g = amd.PPPMSolver(sc.res, sc.dl, sc.dt, sc.min_pos, 0, 0.83, sc.c)
        g.set_mesh(sc.verts, sc.tris)
        return g

    sv = [solver(sc) for sc in scs]
    b = amd.PPPMBatch(sv)
    xs = [torch.from_numpy(sc.neumann).cuda().requires_grad_(True) for sc in scs]
    J = pad.batch_listener_loss(b, sv, xs, targets, lis, segment=13)
    J.sum().backward()
    for i, sc in enumerate(scs):
        s1 = solver(sc)
        x1 = torch.from_numpy(sc.neumann).cuda().requires_grad_(True)
        J1 = pad.listener_loss(x1, targets[i], s1, lis[i], segment=13)
        J1.backward()
        assert abs(float(J1) - float(J[i])) <= 1e-6 * abs(float(J1)), i
        assert torch.equal(x1.grad, xs[i].grad), i
        assert float(x1.grad.abs().max()) > 0


def test_adjoint_identity_on_a_perturbed_l5_scene(amd, scenes_mod):
    import torch
    T = 160
    sc = scenes_mod.c5_scene(0, steps=T, res=64)
    lis = sc.info["listeners"]
    print(f"adjoint identity: {sc.res}^3, {len(sc.tris)} triangles, {len(lis)} listeners, {T} steps", flush=True)
    s = amd.PPPMSolver(sc.res, sc.dl, sc.dt, sc.min_pos, 0, 0.83, sc.c)
    s.set_mesh(sc.verts, sc.tris)
    g = torch.from_numpy(sc.info["target_acc"]).cuda()
    out = torch.zeros(T, len(lis), device="cuda")
    s.run(g, T, lis, out)
    y = torch.from_numpy(np.random.default_rng(4).standard_normal((T, len(lis))).astype(np.float32)).cuda()
    grad = torch.zeros(T, len(sc.tris), device="cuda")
    s.adjoint_reset(T - 1)
    s.adjoint_run(y, T, lis, grad)
    lhs = float((y.double() * out.double()).sum())
    rhs = float((grad.double() * g.double()).sum())
    rel = abs(lhs - rhs) / abs(lhs)
    print(f"  <y, F g> = {lhs:.6e}, <F^T y, g> = {rhs:.6e}, relative difference {rel:.2e}")
    assert abs(lhs) > 0 and rel < 1e-3


def test_descent_iterate_equals_adjoint_of_its_dual_coefficients(amd, scenes_mod):
    """Gradient descent from zero on J_i = |F_i g_i - target_i|^2 keeps g_i = F_i^T w_i with
    w_i = -2 sum_k alpha_k (F_i g_k - target_i): the iterate rebuilt by one batched adjoint sweep
    of w matches the accumulated one (how a long descent continues in another run)."""
    import torch
    import autograd as pad
    T, L = 48, 2
    scs = [scenes_mod.c3_scene(n, lv, seed=n + 3, steps=T) for n, lv in [(24, 1), (30, 2)]]
    lis = [[(3 * sc.res // 4 + 2, sc.res // 2, sc.res // 2), (sc.res // 2, 2 + sc.res // 8, sc.res // 2)]
           for sc in scs]
    print(f"descent dual coefficients: {len(scs)} scenes, {T} steps, {L} listeners", flush=True)
    sv = []
    for sc in scs:
        g = amd.PPPMSolver(sc.res, sc.dl, sc.dt, sc.min_pos, 0, 0.83, sc.c)
        g.set_mesh(sc.verts, sc.tris)
        sv.append(g)
    b = amd.PPPMBatch(sv)
    rng = np.random.default_rng(9)
    targets = [torch.from_numpy(rng.standard_normal((T, L)).astype(np.float32) * 1e-4).cuda() for _ in scs]
    tgt = torch.stack(targets, 1).double()
    gs = [torch.zeros(T, sc.tris.shape[0], device="cuda") for sc in scs]
    w = torch.zeros(T, len(scs), L, device="cuda", dtype=torch.float64)
    smp = torch.zeros(T, len(scs), L, device="cuda")
    out = torch.zeros(T, len(scs), L, device="cuda")
    for _ in range(3):
        xs = [g.detach().requires_grad_(True) for g in gs]
        J = pad.batch_listener_loss(b, sv, xs, targets, lis, segment=7, samples=smp)
        J.sum().backward()
        gr = [x.grad.detach() for x in xs]
        for s in sv:
            s.reset()
        b.run(gr, T, lis, out)
        alphas = [float((gr[i].double() ** 2).sum()) / (2.0 * float((out[:, i].double() ** 2).sum()))
                  for i in range(len(scs))]
        gs = [(xs[i].detach() - alphas[i] * gr[i]).contiguous() for i in range(len(scs))]
        w -= 2.0 * torch.tensor(alphas, device="cuda", dtype=torch.float64)[None, :, None] * (smp.double() - tgt)
    rebuilt = [torch.zeros_like(g) for g in gs]
    b.adjoint_reset(T - 1)
    b.adjoint_run(w.float().contiguous(), T, lis, rebuilt)
    for i in range(len(scs)):
        rel = float((rebuilt[i] - gs[i]).double().norm() / gs[i].double().norm())
        print(f"  scene {i}: |F^T w - g| / |g| = {rel:.2e}")
        assert float(gs[i].abs().max()) > 0 and rel < 1e-4, i

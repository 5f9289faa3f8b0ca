"""Differentiable listener signals of the PPPM step (torch.autograd.Function over the C ABI).

The step (update_grid_and_face, src/algorithm/pppm.h:137-145) is linear in the grid levels, the
history rings and the per-triangle Neumann inputs g (the surface accelerations), so the
gradient of any loss of the listener samples is the adjoint (pppm_adjoint_run: the transposed
step run backwards) applied to the samples' cotangents; it reads no forward state.

ListenerSignal: out = F(g), the listener samples of a run from the zero state; backward is one
adjoint sweep. Memory: the samples (T x n_listen).

ListenerLoss: J = sum (F(g) - target)^2 with sqrt(T) checkpointing: the forward keeps
ceil(T / S) solver states (S = ceil(sqrt T) steps apart) and not the samples; the backward walks
the segments from the last, restores each segment's checkpoint, recomputes that segment's
samples (and so its residuals) and reverses it with the adjoint, whose state carries from one
segment to the one before. The recomputation costs one extra forward run; the gradient equals
ListenerSignal's bit for bit.
"""
from __future__ import annotations

import math

import torch

from pppm_amd import PPPMError

# Checkpoint segments of a single scene shorter than this run with eager launches by default. The
# solver keeps a run's captured graph for later runs whose arguments differ only by the rows of
# one series (pppm_run, pppm_adjoint_run), and the segments below write into full-length
# sample and cotangent buffers so that they do: a segment length is captured once per
# direction, and from a few steps on replaying beats launching.
GRAPH_MIN_SEGMENT = 8
# The same for a batch, whose runs capture their graph per call: a captured segment is replayed
# once, so its capture and instantiation cost more than the launches it saves.
BATCH_GRAPH_MIN_SEGMENT = 256


def _segment_graphs(use_graph, S, minimum=GRAPH_MIN_SEGMENT):
    return S >= minimum if use_graph is None else bool(use_graph)


def _check_inputs(g, solver):
    if not (isinstance(g, torch.Tensor) and g.is_cuda and g.dtype == torch.float32 and g.dim() == 2):
        raise PPPMError("g must be a (steps, triangles) float32 device tensor")
    if g.shape[1] != solver.nt:
        raise PPPMError(f"g has {g.shape[1]} columns, the mesh {solver.nt} triangles")


class ListenerSignal(torch.autograd.Function):
    """out[t, l] = pressure at listeners[l] after step t, from the zero state (solver.reset())."""

    @staticmethod
    def forward(ctx, g, solver, listeners, use_graph=True):
        _check_inputs(g, solver)
        T = g.shape[0]
        solver.reset()
        out = torch.zeros(T, len(listeners), device=g.device, dtype=torch.float32)
        solver.run(g.contiguous(), T, listeners, out, use_graph=use_graph)
        ctx.solver, ctx.listeners, ctx.T, ctx.use_graph = solver, listeners, T, use_graph
        return out

    @staticmethod
    def backward(ctx, gout):
        s = ctx.solver
        grad = torch.zeros(ctx.T, s.nt, device=gout.device, dtype=torch.float32)
        s.adjoint_reset(ctx.T - 1)
        s.adjoint_run(gout.contiguous().float(), ctx.T, ctx.listeners, grad, use_graph=ctx.use_graph)
        return grad, None, None, None


class ListenerLoss(torch.autograd.Function):
    """J = sum_{t, l} (out[t, l] - target[t, l])^2 (FP64 accumulation, returned as float32) with
    sqrt(T) checkpoints; `segment` overrides the checkpoint spacing; use_graph None: graphs for
    segments of GRAPH_MIN_SEGMENT steps or more."""

    @staticmethod
    def forward(ctx, g, target, solver, listeners, segment=None, use_graph=None):
        _check_inputs(g, solver)
        T, L = g.shape[0], len(listeners)
        if tuple(target.shape) != (T, L):
            raise PPPMError("target must be (steps, listeners)")
        S = int(segment) if segment else max(1, math.ceil(math.sqrt(T)))
        nseg = (T + S - 1) // S
        use_graph = _segment_graphs(use_graph, S)
        ck = torch.empty(nseg, solver.state_bytes(), dtype=torch.uint8, device=g.device)
        # full length: every segment's rows sit where its steps are, so segments share a graph
        out = torch.zeros(T, L, device=g.device, dtype=torch.float32)
        J = torch.zeros((), device=g.device, dtype=torch.float64)
        solver.reset()
        for k in range(nseg):
            a, b = k * S, min(T, (k + 1) * S)
            solver.save_state(ck[k])
            solver.run(g[a:b], b - a, listeners, out[a:b], use_graph=use_graph)
            J += ((out[a:b].double() - target[a:b].double()) ** 2).sum()
        ctx.save_for_backward(g, target)
        ctx.solver, ctx.listeners, ctx.S, ctx.nseg, ctx.ck, ctx.use_graph = solver, listeners, S, nseg, ck, use_graph
        ctx.checkpoint_bytes = ck.numel()
        return J.float()

    @staticmethod
    def backward(ctx, gJ):
        g, target = ctx.saved_tensors
        s, S, L = ctx.solver, ctx.S, len(ctx.listeners)
        T = g.shape[0]
        grad = torch.zeros(T, s.nt, device=g.device, dtype=torch.float32)
        out = torch.zeros(T, L, device=g.device, dtype=torch.float32)
        gout = torch.zeros(T, L, device=g.device, dtype=torch.float32)
        s.adjoint_reset(T - 1)
        for k in reversed(range(ctx.nseg)):
            a, b = k * S, min(T, (k + 1) * S)
            s.load_state(ctx.ck[k])
            s.run(g[a:b], b - a, ctx.listeners, out[a:b], use_graph=ctx.use_graph)   # recompute
            torch.mul(out[a:b] - target[a:b].float(), 2.0 * gJ.float(), out=gout[a:b])
            s.adjoint_run(gout[a:b], b - a, ctx.listeners, grad[a:b], use_graph=ctx.use_graph)
        return grad, None, None, None, None, None


def listener_signal(g, solver, listeners, use_graph=True):
    return ListenerSignal.apply(g, solver, listeners, use_graph)


def listener_loss(g, target, solver, listeners, segment=None, use_graph=None):
    return ListenerLoss.apply(g, target, solver, listeners, segment, use_graph)


class BatchListenerLoss(torch.autograd.Function):
    """J_i = sum_{t, l} (out_i[t, l] - target_i[t, l])^2 for every scene i of a batch, with the
    same sqrt(T) checkpointing as ListenerLoss: the forward runs all scenes together
    (pppm_batch_run) in segments of S steps from the zero state and keeps each scene's state at
    the start of every segment; the backward walks the segments from the last, restores every
    scene's checkpoint, recomputes the segment's samples with one batched run and reverses it
    with one batched adjoint run (pppm_batch_adjoint_run: every scene's reverse steps together,
    each scene's gradient bit for bit its own adjoint's). Returns J as an (n_scenes,)
    FP64-accumulated float32 tensor;
    the gradient of scene i depends only on its own inputs (the scenes are independent).
    Inputs: batch (PPPMBatch over `solvers`), targets (one (T, L) tensor per scene), listeners
    (L cells per scene), samples (None, or a (T, n_scenes, L) device tensor that receives the
    listener samples of the forward run), then the per-scene (T, nt_i) inputs g_i."""

    @staticmethod
    def forward(ctx, batch, solvers, targets, listeners, segment, use_graph, samples, *gs):
        n = len(solvers)
        for g, s in zip(gs, solvers):
            _check_inputs(g, s)
        T, L = gs[0].shape[0], len(listeners[0])
        if any(g.shape[0] != T for g in gs) or any(len(x) != L for x in listeners):
            raise PPPMError("every scene needs the same number of steps and of listeners")
        if any(tuple(t.shape) != (T, L) for t in targets):
            raise PPPMError("targets must be (steps, listeners) per scene")
        S = int(segment) if segment else max(1, math.ceil(math.sqrt(T)))
        nseg = (T + S - 1) // S
        use_graph = _segment_graphs(use_graph, S, BATCH_GRAPH_MIN_SEGMENT)
        dev = gs[0].device
        ck = [torch.empty(nseg, s.state_bytes(), dtype=torch.uint8, device=dev) for s in solvers]
        out = torch.zeros(S, n, L, device=dev, dtype=torch.float32)
        J = torch.zeros(n, device=dev, dtype=torch.float64)
        for s in solvers:
            s.reset()
        for k in range(nseg):
            a, b = k * S, min(T, (k + 1) * S)
            for i, s in enumerate(solvers):
                s.save_state(ck[i][k])
            batch.run([g[a:b] for g in gs], b - a, listeners, out[:b - a], use_graph=use_graph)
            if samples is not None:
                samples[a:b] = out[:b - a]
            for i in range(n):
                J[i] += ((out[:b - a, i].double() - targets[i][a:b].double()) ** 2).sum()
        ctx.save_for_backward(*gs)
        ctx.batch, ctx.solvers, ctx.targets, ctx.listeners = batch, solvers, targets, listeners
        ctx.S, ctx.nseg, ctx.ck, ctx.use_graph = S, nseg, ck, use_graph
        ctx.checkpoint_bytes = sum(c.numel() for c in ck)
        return J.float()

    @staticmethod
    def backward(ctx, gJ):
        gs = ctx.saved_tensors
        solvers, S, n = ctx.solvers, ctx.S, len(ctx.solvers)
        T, L = gs[0].shape[0], len(ctx.listeners[0])
        grads = [torch.zeros(T, s.nt, device=g.device, dtype=torch.float32) for g, s in zip(gs, solvers)]
        out = torch.zeros(S, n, L, device=gs[0].device, dtype=torch.float32)
        tgt = torch.stack([t.float() for t in ctx.targets], 1)  # (T, n, L)
        scale = (2.0 * gJ.float())[None, :, None]
        ctx.batch.adjoint_reset(T - 1)
        for k in reversed(range(ctx.nseg)):
            a, b = k * S, min(T, (k + 1) * S)
            for i, s in enumerate(solvers):
                s.load_state(ctx.ck[i][k])
            ctx.batch.run([g[a:b] for g in gs], b - a, ctx.listeners, out[:b - a], use_graph=ctx.use_graph)  # recompute
            gout = scale * (out[:b - a] - tgt[a:b])
            ctx.batch.adjoint_run(gout.contiguous(), b - a, ctx.listeners, [gr[a:b] for gr in grads],
                                  use_graph=ctx.use_graph)
        return (None, None, None, None, None, None, None) + tuple(grads)


def batch_listener_loss(batch, solvers, gs, targets, listeners, segment=None, use_graph=None, samples=None):
    return BatchListenerLoss.apply(batch, solvers, targets, listeners, segment, use_graph, samples, *gs)

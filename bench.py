#!/usr/bin/env python3
"""Benchmark of the PPPM time step on one MI355X.

Default workload (the configuration the metric is quoted on): plain PPPM with motion —
64^3 grid, static subdivided cube (768 tris) + icosphere level 4 (5120 tris) translating
(0.3 h, 200 Hz) and rotating (2 rad/s), band-limited noise on half of the vertices,
4000 steps, FP32 stepping with FP64-built FP16 weights. One "step" = update_mesh (re-bin,
re-list, recompute dirty weight-cache entries) + update_grid_and_face, as the reference's
demo loop does per audio sample (experiments/demo/main.cu:170-205).

Prints one JSON line. value = grid cell-updates/s (N^3 per step); inputs (vertex
positions per step, Neumann series) are resident in HBM before the timed region.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "pppm-tdbem_amd")
sys.path.insert(0, PKG)

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64

# FP64 operations (each +, -, x, /, sqrt counted once) k_grid_weights spends on one dirty
# triangle, counted from its code: 729 layer evaluations (81 targets x 9 wave numbers) of about
# 114 each (distance and its square root, the deterministic complex exponential: exp 33 (range
# reduction + degree-13 Horner) + sin/cos 42, two complex quotients, the double-layer factor),
# 243 stencil batch sums of about 78, the 108 scaled 16-point inverse transforms (4 radix-2 stages
# of 8 per entry, then the scaling: about 59 000) and the tap shifts (about 900).
GRID_WEIGHT_FLOPS_PER_TRIANGLE = 729 * 114 + 243 * 78 + 108 * 16 * (4 * 8 + 2) + 27 * 16 * 2


def algorithmic_bytes(N, st):
    """Algorithmic HBM bytes per launch of each step kernel.

    fdtd      : read p^n, p^{n-1}, write p^{n+1}                  12 B / cell
    far       : per (cell, tri) pair one 72 B FP16 record + 4 B index; per C3 cell read p, write far
    dirichlet : per (tri, tri) pair 72 B record + 4 B index; per triangle 204 B (two 16-tap
                history windows 128 B, 8 interpolation weights 32 B, 8 far reads 32 B,
                area 4 B, N and D writes 8 B)
    near      : as far, reading far and writing p
    """
    P3, C3, P4, T = st["P3"], st["C3"], st["P4"], st["T"]
    return {
        "fdtd": 12 * N ** 3,
        "far": 76 * P3 + 8 * C3,
        "dirichlet": 76 * P4 + 204 * T,
        "near": 76 * P3 + 8 * C3,
    }


def cpu_baseline(sc, steps, threads):
    """The CPU oracle (a port of the reference algorithm; the reference itself is CUDA-only)
    on the host cores: `steps` steps of the same workload after an untimed set_mesh."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    import oracle

    o = oracle.Oracle(sc.res, sc.dl, sc.dt, sc.min_pos, sc.c, 0.83, oracle.FP64)
    o.set_mesh(sc.motion(0) if sc.motion else sc.verts, sc.tris)
    t0 = time.perf_counter()
    for i in range(steps):
        if sc.motion:
            o.update_mesh(sc.motion(i))
        o.step(sc.neumann[i])
    dt = time.perf_counter() - t0
    return sc.res ** 3 * steps / dt, dt


def host_info():
    """nproc and the CPU model of the host the baseline ran on."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count()
    return {"nproc": nproc, "cpu_model": model}


CALIB_MIB = 256


def measure_traffic(kernels, workload_args):
    """HBM bytes per step of the kernels named in `kernels` (one launch each per step; their
    per-dispatch averages are added) from two rocprofv3 counter passes over a short run of
    this script (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each scaled by what the same
    counter reads for a plain 256 MiB device copy in the same run. Returns None if rocprofv3
    is unavailable or a pass fails."""
    import csv
    import shutil
    import tempfile
    if shutil.which("rocprofv3") is None:
        return None
    env = dict(os.environ, TMPDIR="/tmp")
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="pppm_pmc_", dir="/tmp")
        cmd = ["rocprofv3", "--pmc", counter, "--kernel-include-regex", "|".join(kernels) + "|copyBuffer",
               "--output-format", "csv", "-d", d, "-o", "run", "--", sys.executable, os.path.abspath(__file__),
               "--steps", "20", "--warmup", "2", "--timing-steps", "2", "--no-cpu", "--no-traffic",
               "--batch-steps", "0", "--adjoint-steps", "0", "--batched-adjoint-steps", "0",
               "--calib-copy-mib", str(CALIB_MIB)] + workload_args
        try:
            subprocess.run(cmd, env=env, cwd=ROOT, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                           timeout=240, check=True)
            kvals, cvals = {k: [] for k in kernels}, []
            for path, _, files in os.walk(d):
                for f in files:
                    if f.endswith("counter_collection.csv"):
                        for r in csv.DictReader(open(os.path.join(path, f))):
                            if r["Counter_Name"] != counter:
                                continue
                            v = float(r["Counter_Value"])
                            if "copyBuffer" in r["Kernel_Name"]:
                                cvals.append(v)
                                continue
                            for k in kernels:
                                if k + "(" in r["Kernel_Name"] or r["Kernel_Name"].endswith(k):
                                    kvals[k].append(v)
            big = max(cvals) if cvals else 0.0
            if any(not v for v in kvals.values()) or big <= 0:
                return None
            # the copy reads and writes CALIB_MIB MiB
            factor = CALIB_MIB * 1024.0 / big
            res[counter] = {"kib_per_launch": sum(sum(v) / len(v) for v in kvals.values()), "scale": factor}
        except (subprocess.SubprocessError, OSError, KeyError, ValueError):
            return None
        finally:
            shutil.rmtree(d, ignore_errors=True)
    f, w = res["FETCH_SIZE"], res["WRITE_SIZE"]
    return {"bytes": (f["kib_per_launch"] * f["scale"] + w["kib_per_launch"] * w["scale"]) * 1024.0,
            "fetch_kib": f["kib_per_launch"], "fetch_scale": f["scale"],
            "write_kib": w["kib_per_launch"], "write_scale": w["scale"]}


def batch_bench(n_scenes=64, steps=200, warmup=20, timing_steps=20):
    """Batched forward: the 64 uneven scenes (grid edges 24..96, icosphere levels 1..4) stepped
    together by pppm_batch_run (one launch per kernel over every scene, 32 batch steps per graph
    launch, FDTD interior variants picked by the autotuner), against the same scenes run one
    after another through pppm_run (graph-replayed too). Returns the JSON object of the leg."""
    import torch
    import pppm_amd
    import scenes

    sizes = scenes.c3_batch_sizes(n_scenes)
    total = warmup + 2 * steps + timing_steps
    scs = scenes.c3_batch(n_scenes, steps=total)
    t_set = time.perf_counter()
    hs = []
    for sc in scs:
        g = pppm_amd.PPPMSolver(sc.res, sc.dl, sc.dt, sc.min_pos, 0, 0.83, sc.c)
        g.set_mesh(sc.verts, sc.tris)
        hs.append(g)
    torch.cuda.synchronize()
    t_set = time.perf_counter() - t_set
    ns = [torch.from_numpy(sc.neumann).cuda() for sc in scs]
    lis = [sc.listener for sc in scs]
    cells = float(sum(sc.res ** 3 for sc in scs))
    st = [g.stats() for g in hs]
    b = pppm_amd.PPPMBatch(hs)
    t_tune = time.perf_counter()
    table = b.autotune(reps=5)
    t_tune = time.perf_counter() - t_tune
    var = b.variants()
    out = torch.zeros(total, n_scenes, device="cuda")
    i0 = 0

    def rows(a, n):
        return [x[a:a + n] for x in ns]

    b.run(rows(i0, warmup), warmup, lis, out[i0:i0 + warmup])
    i0 += warmup
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.run(rows(i0, steps), steps, lis, out[i0:i0 + steps])
    torch.cuda.synchronize()
    el_b = time.perf_counter() - t0
    i0 += steps
    # per-phase HIP event timing over eager batch steps
    b.set_timing(True)
    b.run(rows(i0, timing_steps), timing_steps, lis, out[i0:i0 + timing_steps], use_graph=False)
    b.set_timing(False)
    ph = b.get_timing()
    i0 += timing_steps
    # the same scenes one after another (pppm_run, graph-replayed), same number of steps
    o1 = torch.zeros(steps, 1, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for g, x, l in zip(hs, ns, lis):
        g.run(x[i0:i0 + steps], steps, [l], o1)
    torch.cuda.synchronize()
    el_s = time.perf_counter() - t0

    by = {"fdtd": 0.0, "windows": 0.0, "far": 0.0, "dirichlet": 0.0, "near": 0.0}
    for sc, s_ in zip(scs, st):
        a = algorithmic_bytes(sc.res, s_)
        for k in ("fdtd", "far", "dirichlet", "near"):
            by[k] += a[k]
        by["windows"] += 256.0 * s_["T"]  # two 16-value ring windows read and written per triangle
    avg = {k: (v[0] / v[1] if v[1] else None) for k, v in ph.items()}
    step_kernels = ["fdtd", "far", "dirichlet", "near"]
    dom = max(step_kernels, key=lambda k: avg[k] or 0.0)
    ach = by[dom] / (avg[dom] * 1e-3) / 1e9
    step_bytes = sum(by[k] for k in step_kernels)
    kname = {"fdtd": "k_batch_fdtd + k_batch_tile", "far": "k_batch_far", "dirichlet": "k_batch_dirichlet",
             "near": "k_batch_near"}[dom]
    return {
        "metric": "batched grid cell-updates/s",
        "value": cells * steps / el_b,
        "unit": "cell-updates/s",
        "steps": steps,
        "ms_per_batch_step": el_b * 1e3 / steps,
        "sequential_value": cells * steps / el_s,
        "sequential_ms_per_step_all_scenes": el_s * 1e3 / steps,
        "speedup_vs_sequential": el_s / el_b,
        "config": {"workload": f"c3_batch_{n_scenes}_scenes", "scenes": n_scenes, "grid_edges": [r for r, _ in sizes],
                   "levels": [v for _, v in sizes], "cells_per_step": int(cells),
                   "triangles": int(sum(s_["T"] for s_ in st)), "p3_pairs": int(sum(s_["P3"] for s_ in st)),
                   "p4_pairs": int(sum(s_["P4"] for s_ in st)), "c3_cells": int(sum(s_["C3"] for s_ in st))},
        "algorithmic_bytes_per_batch_step": step_bytes,
        "hbm_gbps_effective_step": step_bytes / (el_b / steps) / 1e9,
        "hbm_frac_effective_step": step_bytes / (el_b / steps) / 1e9 / HBM_PEAK_GBPS,
        "roofline": {"kernel": kname, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBPS, "bytes_per_launch": by[dom], "avg_launch_us": avg[dom] * 1e3},
        "phase_avg_us": {k: (v * 1e3 if v else None) for k, v in avg.items()},
        "phase_gbps": {k: (by[k] / (avg[k] * 1e-3) / 1e9 if avg[k] else None) for k in by},
        "autotune_ms": {str(k): v for k, v in table.items()},
        "variants_used": {pppm_amd.PPPMBatch.VARIANTS[v]: int((var == v).sum()) for v in sorted(set(var.tolist()))},
        "setup_s": {"set_mesh_all_scenes": t_set, "autotune": t_tune},
    }


def adjoint_bench(steps=2000):
    """Adjoint: the 48^3 / icosphere L3 scene, `steps` steps. Times a forward run (pppm_run),
    the adjoint sweep over the same steps (pppm_adjoint_run), and the sqrt(T)-checkpointed
    listener loss with its backward (forward + recompute + adjoint). Returns the leg's object."""
    import torch
    import pppm_amd
    import scenes
    import autograd as pad

    sc = scenes.c4_adjoint(steps=steps)
    s = pppm_amd.PPPMSolver(sc.res, sc.dl, sc.dt, sc.min_pos, 0, 0.83, sc.c)
    s.set_mesh(sc.verts, sc.tris)
    g = torch.from_numpy(sc.neumann).cuda()
    lis = [sc.listener]
    out = torch.zeros(steps, 1, device="cuda")
    gout = torch.randn(steps, 1, device="cuda")
    grad = torch.zeros(steps, len(sc.tris), device="cuda")
    # warm-up of both directions (set-up of the transposed lists happens in the first reset)
    s.reset()
    s.run(g[:64], 64, lis, out[:64])
    s.adjoint_reset(63)
    s.adjoint_run(gout[:64], 64, lis, grad[:64])
    torch.cuda.synchronize()
    s.reset()
    t0 = time.perf_counter()
    s.run(g, steps, lis, out)
    torch.cuda.synchronize()
    el_f = time.perf_counter() - t0
    s.adjoint_reset(steps - 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.adjoint_run(gout, steps, lis, grad)
    torch.cuda.synchronize()
    el_a = time.perf_counter() - t0
    target = torch.zeros(steps, 1, device="cuda")
    gg = g.clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J = pad.listener_loss(gg, target, s, lis)
    J.backward()
    torch.cuda.synchronize()
    el_l = time.perf_counter() - t0
    N = sc.res
    return {
        "metric": "adjoint grid cell-updates/s",
        "value": N ** 3 * steps / el_a,
        "unit": "cell-updates/s",
        "steps": steps,
        "forward_ms_per_step": el_f * 1e3 / steps,
        "adjoint_ms_per_step": el_a * 1e3 / steps,
        "adjoint_over_forward": el_a / el_f,
        "checkpointed_loss_and_gradient_over_forward": el_l / el_f,
        "checkpoint_segments": int(math.ceil(steps / math.ceil(math.sqrt(steps)))),
        "config": {"workload": "c4_adjoint", "grid": f"{N}^3", "triangles": len(sc.tris),
                   "gradient_shape": [steps, len(sc.tris)]},
    }


def c5_bench(n_scenes=16, steps=8192, iters=3, res=128, seed_dir=11, step_rule="exact", save_dual=None,
             resume=None):
    """Batched adjoint at scale: n_scenes scenes of res^3 with a radially perturbed icosphere L5
    (20480 triangles) and 8 listeners each, `steps` steps, source reconstruction by gradient
    descent on the per-triangle accelerations: J_i = sum (F_i(g_i) - target_i)^2, target_i the
    listener signals of seeded accelerations, g_i from zero. Each iteration: the batched,
    sqrt(T)-checkpointed loss and its gradient (autograd.BatchListenerLoss: batched forward,
    batched recompute and batched adjoint per segment), then a step along -grad: step_rule
    "exact" takes the exact line search (J is quadratic: alpha_i = |grad_i|^2 / (2 |F_i grad_i|^2),
    one batched forward); "bb" takes it on the first iteration and the Barzilai-Borwein step
    alpha_i = s.s / s.y after it (s, y: the changes of the iterate and of the gradient over the
    last iteration; s.y = 2 |F s|^2 > 0 for this J), which needs no extra forward.
    A descent can span several runs: every gradient is F^T (2 r) (r = the listener residuals, a
    (steps, L) series per scene), so the iterate from zero is g = F^T w with w = -2 sum alpha r.
    save_dual writes w and the loss history (a few MB); resume reads them and rebuilds g with one
    batched adjoint sweep (its first iteration then takes the exact line search).
    Spot checks on scene 0 at full length: the adjoint identity <y, F g> = <F^T y, g> for random
    g, y, and the directional derivative (J(g + d) - J(g - d)) / 2 against <grad J, d>, which J
    being quadratic makes exact up to rounding. Returns the leg's JSON object."""
    import torch
    import pppm_amd
    import scenes
    import autograd as pad

    def note(msg):  # progress on stderr (the JSON line stays alone on stdout)
        print(f"[c5 {time.perf_counter() - t_start:7.1f} s] {msg}", file=sys.stderr, flush=True)

    t_start = time.perf_counter()
    free0, total = torch.cuda.mem_get_info()
    scs = []
    for i in range(n_scenes):
        scs.append(scenes.c5_scene(i, steps=steps, res=res))
        note(f"scene {i + 1}/{n_scenes} generated")
    t_gen = time.perf_counter() - t_start
    t0 = time.perf_counter()
    sv = []
    for sc in scs:
        g = pppm_amd.PPPMSolver(sc.res, sc.dl, sc.dt, sc.min_pos, 0, 0.83, sc.c)
        g.set_mesh(sc.verts, sc.tris)
        sv.append(g)
    torch.cuda.synchronize()
    t_setmesh = time.perf_counter() - t0
    note(f"set_mesh of {n_scenes} scenes: {t_setmesh:.1f} s")
    batch = pppm_amd.PPPMBatch(sv)
    lis = [sc.info["listeners"] for sc in scs]
    L, nt = len(lis[0]), len(scs[0].tris)
    dev = "cuda"
    # targets: listener signals of the seeded accelerations
    true = [torch.from_numpy(sc.info["target_acc"]).to(dev) for sc in scs]
    for sc in scs:
        sc.info["target_acc"] = None
    out = torch.zeros(steps, n_scenes, L, device=dev)
    for s_ in sv:
        s_.reset()
    batch.run(true, steps, lis, out)
    targets = [out[:, i].clone() for i in range(n_scenes)]
    note("targets computed")
    # plain batched forward and batched adjoint sweep over the whole length (no checkpoints)
    for s_ in sv:
        s_.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch.run(true, steps, lis, out)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    note(f"batched forward: {t_fwd:.1f} s")
    fwd_out0 = out[:, 0].clone()
    gen = torch.Generator(device=dev).manual_seed(seed_dir)
    # cotangents y_i = F_i g_i (so <y, F g> = |F g|^2 has no cancellation)
    y = [out[:, i].clone() for i in range(n_scenes)]
    grads = [torch.zeros(steps, nt, device=dev) for _ in range(n_scenes)]
    batch.adjoint_reset(steps - 1)  # the transposed lists are built once per mesh, outside the timing
    yb = torch.stack(y, 1).contiguous()  # (steps, n_scenes, L): the layout of the batched forward's output
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch.adjoint_reset(steps - 1)
    batch.adjoint_run(yb, steps, lis, grads)
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    del yb
    note(f"batched adjoint sweep: {t_adj:.1f} s")
    # adjoint identity on scene 0: <y, F g> = <F^T y, g> with g = the seeded accelerations
    lhs = float((y[0].double() * fwd_out0.double()).sum())
    rhs = float((grads[0].double() * true[0].double()).sum())
    dot_rel = abs(lhs - rhs) / max(abs(lhs), 1e-300)
    del true, grads, y
    # gradient descent (from zero, or from the iterate F^T w of an earlier run)
    gs = [torch.zeros(steps, nt, device=dev) for _ in range(n_scenes)]
    w_dual = torch.zeros(steps, n_scenes, L, device=dev, dtype=torch.float64)
    hist, it0, t_resume = [], 0, 0.0
    if resume:
        z = np.load(resume)
        if tuple(z["w"].shape) != (steps, n_scenes, L) or int(z["res"]) != res:
            raise SystemExit(f"{resume}: saved for another configuration")
        w_dual = torch.from_numpy(z["w"]).to(dev)
        hist = [h for h in z["hist"]]
        it0 = int(z["iterations"])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        batch.adjoint_reset(steps - 1)
        batch.adjoint_run(w_dual.float().contiguous(), steps, lis, gs)
        torch.cuda.synchronize()
        t_resume = time.perf_counter() - t0
        note(f"resumed after {it0} iterations: iterate rebuilt from its dual coefficients in {t_resume:.1f} s")
    samples = torch.zeros(steps, n_scenes, L, device=dev)
    tgt = torch.stack(targets, 1).double()
    prev_x = prev_g = None
    t_loss, t_back, t_ls = 0.0, 0.0, 0.0
    ck_bytes = 0
    torch.cuda.reset_peak_memory_stats()
    for it in range(iters):
        xs = [g.detach().requires_grad_(True) for g in gs]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        J = pad.batch_listener_loss(batch, sv, xs, targets, lis, samples=samples)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        J.sum().backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        t_loss += t1 - t0
        t_back += t2 - t1
        ck_bytes = sum(sv_.state_bytes() for sv_ in sv) * math.ceil(steps / math.ceil(math.sqrt(steps)))
        hist.append(J.detach().double().cpu().numpy())
        note(f"iteration {it0 + it + 1}/{it0 + iters}: loss / initial {float((hist[-1] / hist[0]).mean()):.4f}; "
             f"forward {t1 - t0:.1f} s, recompute + adjoint {t2 - t1:.1f} s")
        gr = [x.grad.detach() for x in xs]
        alphas = []
        if step_rule == "bb" and prev_x is not None:
            # Barzilai-Borwein: alpha_i = s.s / s.y
            for i in range(n_scenes):
                sd = (xs[i].detach() - prev_x[i]).double()
                yd = (gr[i] - prev_g[i]).double()
                sy = float((sd * yd).sum())
                alphas.append(float((sd * sd).sum()) / sy if sy > 0 else 0.0)
                del sd, yd
        else:
            # exact line search along -grad: alpha_i = |grad_i|^2 / (2 |F_i grad_i|^2)
            for s_ in sv:
                s_.reset()
            batch.run(gr, steps, lis, out)
            for i in range(n_scenes):
                den = 2.0 * float((out[:, i].double() ** 2).sum())
                alphas.append(float((gr[i].double() ** 2).sum()) / den if den > 0 else 0.0)
        torch.cuda.synchronize()
        t_ls += time.perf_counter() - t2
        if step_rule == "bb":
            prev_x, prev_g = [x.detach() for x in xs], gr
        for i in range(n_scenes):
            gs[i] = (xs[i].detach() - alphas[i] * gr[i]).contiguous()
        w_dual -= 2.0 * torch.tensor(alphas, device=dev, dtype=torch.float64)[None, :, None] * (samples.double() - tgt)
        del xs, gr
    # directional derivative on scene 0 at the current iterate (any point works for a quadratic),
    # along a random direction with a component along the gradient (no cancellation in <grad, d>)
    g_prev = gs[0]

    def J0(g):
        o = torch.zeros(steps, L, device=dev)
        sv[0].reset()
        sv[0].run(g.contiguous(), steps, lis[0], o)
        return float(((o.double() - targets[0].double()) ** 2).sum())

    xs0 = g_prev.detach().requires_grad_(True)
    Jc = pad.listener_loss(xs0, targets[0], sv[0], lis[0])
    Jc.backward()
    gn = xs0.grad.detach()
    d = torch.randn(steps, nt, device=dev, generator=gen) * float(gn.abs().mean()) - gn
    d = d / float(d.abs().max())
    fd = (J0(g_prev + d) - J0(g_prev - d)) / 2.0
    an = float((xs0.grad.double() * d.double()).sum())
    dir_rel = abs(fd - an) / max(abs(an), 1e-300)
    free1, _ = torch.cuda.mem_get_info()
    hist = np.array(hist)
    if save_dual:
        np.savez(save_dual, w=w_dual.cpu().numpy(), hist=hist, iterations=it0 + iters, res=res)
        note(f"dual coefficients after {it0 + iters} iterations written to {save_dual}")
    cells = n_scenes * res ** 3 * steps
    per_it = (t_loss + t_back + t_ls) / iters
    res_ = {
        "metric": "batched adjoint: source reconstruction by gradient descent",
        "config": {"workload": f"c5_batched_adjoint_{n_scenes}x{res}^3", "scenes": n_scenes, "grid": f"{res}^3",
                   "triangles_per_scene": nt, "listeners_per_scene": L, "steps": steps, "iterations": iters,
                   "iterations_before": it0, "step_rule": step_rule},
        "forward_cell_updates_per_s": cells / t_fwd,
        "adjoint_cell_updates_per_s": cells / t_adj,
        "forward_s": t_fwd, "adjoint_s": t_adj, "adjoint_over_forward": t_adj / t_fwd,
        "iteration_s": per_it,
        "iteration_breakdown_s": {"checkpointed_forward": t_loss / iters, "recompute_and_adjoint": t_back / iters,
                                  "step_size": t_ls / iters},
        "checkpoint_recompute_overhead": (t_loss + t_back) / iters / (t_fwd + t_adj) - 1.0,
        "checkpoint_bytes": ck_bytes,
        "memory_gb": {"device_used": (total - free1) / 1e9, "device_used_before": (total - free0) / 1e9,
                      "torch_peak_during_descent": torch.cuda.max_memory_allocated() / 1e9},
        "loss_relative_to_initial": (hist[:, :] / hist[0]).mean(axis=1).tolist(),
        "loss_per_scene_first_last": [hist[0].tolist()[:4], hist[-1].tolist()[:4]],
        "spot_checks": {"adjoint_identity_rel": dot_rel, "directional_derivative_rel": dir_rel},
        "setup_s": {"scene_generation": t_gen, "set_mesh_all": t_setmesh, "resume_adjoint_sweep": t_resume},
        "projected_s_for_50_iterations": 50 * per_it,
    }
    return res_


def modal_bench(nt=20480, nmodes=200, nsamples=4096, calls=5, cpu=True):
    """Modal sound source (surface accelerations of a rigid body) on one MI355X: nt surface
    triangles (icosphere level 5 size), nmodes modes, blocks of nsamples audio samples per call.
    Prints one JSON line; the step-path metric is the default run."""
    import torch
    import pppm_amd
    rng = np.random.default_rng(11)
    nv = nt // 2 + 2
    tris = rng.integers(0, nv, (nt, 3)).astype(np.int32)
    n = rng.standard_normal((nt, 3))
    n = (n / np.linalg.norm(n, axis=1, keepdims=True)).astype(np.float32)
    eig = (rng.standard_normal((3 * nv, nmodes)) * 0.1).astype(np.float32)
    eigvals = ((2 * np.pi * np.sort(rng.uniform(100.0, 15000.0, nmodes))) ** 2).astype(np.float32)
    ms = pppm_amd.ModalSource(eig, eigvals, tris, n, 30.0, 1e-6, 44100.0)
    forces = torch.from_numpy((rng.standard_normal((nsamples, nmodes)) * 1e3).astype(np.float32)).cuda()
    accs = torch.empty((nsamples, nt), dtype=torch.float32, device="cuda")
    ms.run(forces, accs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(calls):
        ms.run(forces, accs)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / calls
    flops = 2.0 * nt * nmodes * nsamples
    bytes_alg = 4.0 * (nt * nmodes + nsamples * nmodes + nsamples * nt)
    res = {"metric": "modal surface accelerations/s", "value": nt * nsamples / el, "unit": "triangle-samples/s",
           "n_gpus": 1, "ms_per_call": el * 1e3, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "modal_source", "triangles": nt, "modes": nmodes, "samples_per_call": nsamples},
           "roofline": {"bound": "valu-fp32", "achieved": flops / el / 1e12, "unit": "TFLOP/s",
                        "peak": 157.3, "frac": flops / el / 1e12 / 157.3,
                        "note": "whole call (oscillators + accelerations); sequential FP32 sums in mode order, "
                                "one fused multiply-add per term, no packed FP32 (ceiling 78.6 TFLOP/s); "
                                "algorithmic bytes %.0f MB -> %.0f GB/s" % (bytes_alg / 1e6, bytes_alg / el / 1e9)}}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        c = oracle.modal_coeffs(eigvals, 1.0 / 44100.0, 30.0, 1e-6)
        m = oracle.modal_surface_matrix(eig, tris, n)
        state = np.zeros((nmodes, 3), np.float32)
        f_all = forces.cpu().numpy()
        ns, dt = 0, 0.0
        while ns < nsamples and dt < 10.0:  # about 10 s of CPU work, in blocks of 256 samples
            f = f_all[ns:ns + 256]
            t0 = time.perf_counter()
            q = oracle.modal_iir(c, 1.0 / 44100.0, state, f)
            oracle.modal_accs(m, q)
            dt += time.perf_counter() - t0
            ns += f.shape[0]
        res["cpu_baseline"] = {"value": nt * ns / dt, "unit": "triangle-samples/s", "cores": 1, "kind": "port",
                               "sample": f"oracle IIR + accelerations, {ns} samples, {dt:.2f} s"}
    print(json.dumps(res))


def ghost_bench(res=64, steps=500, warmup=20, cpu=True):
    """Ghost-cell comparison solver (GhostCellSolver::update: FDTD step, boundary data, ghost
    solve, SOLID cells to zero) on one MI355X: the C1 sphere on a res^3 grid, condition
    threshold 25. Prints one JSON line; the step-path metric is the default run."""
    import torch
    import pppm_amd
    import scenes
    sc = scenes.c1_static_sphere(steps=warmup + steps, res=res)
    g = pppm_amd.PPPMSolver(sc.res, sc.dl, sc.dt, sc.min_pos, 0, 0.83, sc.c)
    g.set_mesh(sc.verts, sc.tris)
    n = g.classify(False)
    second = g.ghost_setup(25.0)
    neu = torch.from_numpy(sc.neumann).cuda()
    for i in range(warmup):
        g.ghost_step(neu[i])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(warmup, warmup + steps):
        g.ghost_step(neu[i])
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    it = g.ghost_solve()
    res_ = {"metric": "ghost-cell solver cell-updates/s", "value": res ** 3 * steps / el, "unit": "cell-updates/s",
            "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": el * 1e3 / steps, "higher_is_better": True,
            "dtype": "f32 (solve f64)", "data": "synthetic",
            "config": {"workload": f"c1_static_sphere_{res}_ghost_cell", "grid": f"{res}^3", "ghost_cells": n,
                       "second_order_rows": second, "solver_iterations_last_step": it}}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        import ghost_oracle
        o = oracle.Oracle(sc.res, sc.dl, sc.dt, sc.min_pos, sc.c, 0.83, oracle.FP64)
        cells = g.cells()
        st = ghost_oracle.ghost_setup(cells, sc.res, sc.dl, sc.min_pos, 25.0)
        k, dt = 0, 0.0
        while k < steps and dt < 10.0:
            t1 = time.perf_counter()
            ghost_oracle.ghost_step(o, st, cells, sc.neumann[k])
            dt += time.perf_counter() - t1
            k += 1
        res_["cpu_baseline"] = {"value": res ** 3 * k / dt, "unit": "cell-updates/s", "cores": 1, "kind": "port",
                                "sample": f"oracle FDTD + numpy ghost solve (dense LU), {k} steps, {dt:.1f} s"}
    print(json.dumps(res_))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: the workload's 4000)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--c5-scenes", type=int, default=16)
    ap.add_argument("--c5-steps", type=int, default=8192)
    ap.add_argument("--c5-iters", type=int, default=3)
    ap.add_argument("--c5-res", type=int, default=128)
    ap.add_argument("--c5-save-dual", default=None,
                    help="write the descent's dual coefficients and loss history here (.npz) at the end")
    ap.add_argument("--c5-resume", default=None, help="continue the descent saved by --c5-save-dual")
    ap.add_argument("--c5-step", default="exact", choices=["exact", "bb"],
                    help="gradient-descent step: exact line search (one batched forward per iteration) or "
                         "Barzilai-Borwein after the first iteration")
    ap.add_argument("--workload", default="c2_moving",
                    choices=["c2_moving", "c1_static_sphere", "large_static_sphere", "c5_batched_adjoint"],
                    help="c2_moving: the metric's workload; c1_static_sphere: 32^3 static sphere; "
                         "large_static_sphere: one 512^3 grid around the C5 object (HBM-resident levels)")
    ap.add_argument("--timing-steps", type=int, default=200, help="extra steps timed per kernel with HIP events")
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--static-geometry", action="store_true",
                    help="profiling variant: the moving workload's first frame held static, steps replayed "
                         "through the captured graph (step kernels only, no mesh updates)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE) behind roofline.traffic")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="after the run, copy a buffer of this many MiB five times (a plain copy kernel of "
                         "known size, to scale FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--modal", action="store_true",
                    help="measure the modal sound source instead (its own JSON line)")
    ap.add_argument("--ghost", action="store_true",
                    help="measure the ghost-cell comparison solver instead (its own JSON line)")
    ap.add_argument("--batch-steps", type=int, default=200,
                    help="timed batch steps of the 64-scene batched leg (0: leg skipped)")
    ap.add_argument("--adjoint-steps", type=int, default=2000,
                    help="steps of the adjoint leg (48^3, icosphere L3; 0: leg skipped)")
    ap.add_argument("--batched-adjoint-steps", type=int, default=256,
                    help="steps of the reduced batched-adjoint leg (4 x 64^3, perturbed icosphere L5, "
                         "3 gradient-descent iterations; 0: leg skipped)")
    args = ap.parse_args()
    if args.modal:
        modal_bench(cpu=not args.no_cpu)
        return
    if args.ghost:
        ghost_bench(cpu=not args.no_cpu)
        return
    if args.workload == "c5_batched_adjoint":
        print(json.dumps(c5_bench(args.c5_scenes, args.c5_steps, args.c5_iters, args.c5_res, step_rule=args.c5_step,
                                  save_dual=args.c5_save_dual, resume=args.c5_resume)))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is a design note, "
                                   "not an implemented path", "n_gpus_requested": args.gpus}))
        sys.exit(2)

    import torch
    import pppm_amd
    import scenes

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU"}))
        sys.exit(1)

    default_steps = {"c2_moving": 4000, "c1_static_sphere": 1000, "large_static_sphere": 500}[args.workload]
    K = args.steps or default_steps
    W = args.warmup
    total = W + K + args.timing_steps
    sc = {"c2_moving": scenes.c2_moving, "c1_static_sphere": scenes.c1_static_sphere,
          "large_static_sphere": scenes.large_static_sphere}[args.workload](steps=total)
    if args.static_geometry:
        sc.motion = None
        sc.name = sc.name + "_static_geometry"
    moving = sc.motion is not None
    N = sc.res

    # inputs resident in HBM
    neumann = torch.from_numpy(sc.neumann).cuda()
    if moving:
        verts_all = torch.from_numpy(np.stack([sc.motion(i) for i in range(total)])).cuda()
    torch.cuda.synchronize()

    stream = torch.cuda.Stream()
    solver = pppm_amd.PPPMSolver(sc.res, sc.dl, sc.dt, sc.min_pos, 0, 0.83, sc.c)
    solver.set_stream(stream.cuda_stream)
    solver.set_mesh(verts_all[0] if moving else sc.verts, torch.from_numpy(sc.tris).cuda() if moving else sc.tris)
    lis = [sc.listener]
    out = torch.zeros(total, 1, device="cuda")

    def run_steps(i0, n, use_graph=True):
        if moving:
            # the demo loop's per-sample UpdateSound: update_mesh, then update_grid_and_face
            # (experiments/demo/main.cu:170-205), for every step of the window in one call
            # (pppm_run_moving: the update's status is read while the step runs); the vertex and
            # Neumann rows are already in HBM
            solver.run_moving(verts_all[i0:i0 + n], neumann[i0:i0 + n], n, lis, out[i0:i0 + n])
        else:
            solver.run(neumann[i0:i0 + n], n, lis, out[i0:i0 + n], use_graph=use_graph)

    with torch.cuda.stream(stream):
        run_steps(0, W)
        stream.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run_steps(W, K)
        stream.synchronize()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        elapsed = t1 - t0
        st = solver.stats()
        # per-kernel durations: HIP events around each launch, on the solver's stream
        solver.set_timing(True)
        run_steps(W + K, args.timing_steps, use_graph=False)
        solver.set_timing(False)
        phases = solver.get_timing()
        work = solver.get_update_work()
        stream.synchronize()

    # the practical HBM ceiling beside the published peak: a plain 1 GiB device-to-device copy
    # (read + write bytes), timed with HIP events; not in the counter passes' runs, whose only
    # large copies must be the calibration copies
    copy_gbps = None
    if args.calib_copy_mib == 0:
        try:
            ca = torch.ones(1 << 28, device="cuda")
            cb = torch.empty_like(ca)
            cb.copy_(ca)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(10):
                cb.copy_(ca)
            e1.record()
            torch.cuda.synchronize()
            copy_gbps = 2 * ca.numel() * 4 * 10 / (e0.elapsed_time(e1) * 1e-3) / 1e9
            del ca, cb
        except RuntimeError:  # no memory for the two 1 GiB buffers
            pass
    if args.calib_copy_mib > 0:
        n = args.calib_copy_mib * (1 << 20) // 4
        a = torch.ones(n, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    listener = float(solver.read_cells(lis)[0])
    ms_step = elapsed * 1e3 / K
    value = N ** 3 * K / elapsed
    bytes_ = algorithmic_bytes(N, st)
    step_bytes = sum(bytes_.values())
    avg = {k: (phases[k][0] / phases[k][1] if phases[k][1] else None) for k in phases}
    step_kernels = ["fdtd", "far", "dirichlet", "near"]
    dom = max(step_kernels, key=lambda k: avg[k] or 0.0)
    achieved = bytes_[dom] / (avg[dom] * 1e-3) / 1e9
    # shares of the timed step: the update's own phases (grid weights, face cache, lists) are
    # parts of update_mesh, so the step total is fdtd + far + dirichlet + near + update_mesh
    top = ["fdtd", "far", "dirichlet", "near", "update_mesh"]
    step_total = max(sum(phases[k][0] for k in top), 1e-12)
    phase_share = {k: round(phases[k][0] / step_total, 4) for k in phases}
    # per-kernel rooflines of every phase above 10 % of the step: HBM for the step kernels
    # (algorithmic bytes), FP64 for the grid-cache recompute (operations per dirty triangle)
    kernels = {}
    for k, kn in (("fdtd", "k_fdtd_abc"), ("far", "k_far"), ("dirichlet", "k_dirichlet"), ("near", "k_near")):
        if avg[k]:
            a = bytes_[k] / (avg[k] * 1e-3) / 1e9
            kernels[kn] = {"share_of_step": phase_share[k], "avg_us": avg[k] * 1e3, "bound": "hbm",
                           "bytes_per_launch": bytes_[k], "achieved": a, "unit": "GB/s", "frac": a / HBM_PEAK_GBPS}
    if avg.get("grid_weights") and work["updates"]:
        tri = work["grid_triangles"] / work["updates"]
        fl = tri * GRID_WEIGHT_FLOPS_PER_TRIANGLE
        a = fl / (avg["grid_weights"] * 1e-3) / 1e12
        kernels["k_grid_weights"] = {"share_of_step": phase_share["grid_weights"], "avg_us": avg["grid_weights"] * 1e3,
                                     "bound": "fp64", "dirty_triangles_per_update": tri,
                                     "fp64_ops_per_triangle": GRID_WEIGHT_FLOPS_PER_TRIANGLE, "achieved": a,
                                     "unit": "TFLOP/s", "peak": FP64_PEAK_TFLOPS, "frac": a / FP64_PEAK_TFLOPS}
    for k, kn in (("face_cache", "face-cache update (k_face_lookup, scan, k_face_regular/_singular/_finish)"),
                  ("lists", "geometry, binning and lists (k_tri_update, scans, k_lists_count/_fill)")):
        if avg.get(k):
            kernels[kn] = {"share_of_step": phase_share[k], "avg_us": avg[k] * 1e3}
    if work["updates"]:
        kernels["update_work"] = {"grid_triangles_per_update": work["grid_triangles"] / work["updates"],
                                  "face_pairs_per_update": work["face_pairs"] / work["updates"]}

    res = {
        "metric": "grid cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": sc.name,
            "grid": f"{N}^3",
            "triangles": int(st["T"]),
            "moving": moving,
            "c3_cells": st["C3"], "p3_pairs": st["P3"], "p4_pairs": st["P4"],
            "dt": sc.dt, "h": sc.dl,
        },
        "hbm_gbps_effective_step_kernels": step_bytes / (sum(avg[k] for k in step_kernels) * 1e-3) / 1e9,
        "roofline": {
            # from an interior edge of 128 cells the FDTD step is two launches (marched interior, ABC shell)
            "kernel": {"fdtd": "k_fdtd_march + k_fdtd_abc" if N - 4 >= 128 else "k_fdtd_abc", "far": "k_far",
                       "dirichlet": "k_dirichlet", "near": "k_near"}[dom],
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": None,
            "bytes_per_launch": bytes_[dom],
            "avg_launch_us": avg[dom] * 1e3,
            "device_copy_gbps": copy_gbps,
        },
        "hbm_gbps_effective_step": step_bytes / (ms_step * 1e-3) / 1e9,
        "phase_avg_us": {k: (v * 1e3 if v else None) for k, v in avg.items()},
        "phase_time_share": phase_share,
        "kernels": kernels,
        "listener_final": listener,
        # face-cache slots handed out since set_mesh (never reclaimed) against the live face pairs
        "face_cache": {"slots_used": st["pool"], "live_pairs": st["P4"]},
    }
    if not args.no_traffic:
        kname = res["roofline"]["kernel"]
        wl = ["--workload", args.workload] + (["--static-geometry"] if args.static_geometry else [])
        tr = measure_traffic(kname.split(" + "), wl)
        if tr is not None:
            res["roofline"]["traffic"] = tr["bytes"]
            res["roofline"]["traffic_detail"] = {
                "source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, 20-step run of this workload",
                "fetch_kib_per_launch": tr["fetch_kib"], "fetch_scale": tr["fetch_scale"],
                "write_kib_per_launch": tr["write_kib"], "write_scale": tr["write_scale"],
                "calibration": f"{CALIB_MIB} MiB device copy"}
    if args.batch_steps > 0:
        res["batched"] = batch_bench(steps=args.batch_steps)
    if args.adjoint_steps > 0:
        res["adjoint"] = adjoint_bench(steps=args.adjoint_steps)
    if args.batched_adjoint_steps > 0:
        # the batched adjoint at a reduced size (the full 16 x 128^3 x 8192-step workload runs with
        # --workload c5_batched_adjoint)
        res["batched_adjoint"] = c5_bench(n_scenes=4, steps=args.batched_adjoint_steps, iters=3, res=64)
    if not args.no_cpu:
        cpu_rate, cpu_s = cpu_baseline(sc, args.cpu_steps, args.cpu_threads)
        one_steps = max(1, args.cpu_steps // 4)
        one_rate, one_s = cpu_baseline(sc, one_steps, 1)
        res["cpu_baseline"] = {
            "value": cpu_rate,
            "unit": "cell-updates/s",
            "cores": args.cpu_threads,
            "kind": "port",
            "sample": f"CPU oracle (OpenMP, FP64 weights), first {args.cpu_steps} steps of {sc.name} incl. "
                      f"update_mesh, after an untimed set_mesh; {cpu_s:.1f} s",
            "value_1_thread": one_rate,
            "sample_1_thread": f"first {one_steps} steps, 1 thread; {one_s:.1f} s",
            "host": host_info(),
        }
    print(json.dumps(res))


if __name__ == "__main__":
    main()

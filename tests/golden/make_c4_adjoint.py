"""Writes tests/golden/c4_adjoint.npz: the oracle's FP64 reference for the adjoint configuration
(48^3, icosphere L3, 2000 steps; scenes.c4_adjoint): the listener target, the loss at the
inputs, the FP64 adjoint gradient at 4096 sampled (step, triangle) entries plus its norm, and
central finite differences (eps = 1e-3, FP64 forward runs) at 32 random entries.
Run from the repository root: python tests/golden/make_c4_adjoint.py"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "pppm-tdbem_amd"), os.path.join(ROOT, "oracle")]
import oracle  # noqa: E402
import scenes  # noqa: E402


def main():
    t0 = time.time()
    sc = scenes.c4_adjoint()
    T, nt = sc.steps, len(sc.tris)
    o = oracle.Oracle(sc.res, sc.dl, sc.dt, sc.min_pos, sc.c, 0.83, oracle.FP64)
    o.set_mesh(sc.verts, sc.tris)
    lis = [sc.listener]
    target = o.linear_forward(sc.info["target_acc"], lis)
    g = sc.neumann.astype(np.float64)
    out = o.linear_forward(g, lis)
    loss = float(np.sum((out - target) ** 2))
    grad = o.linear_adjoint(2 * (out - target), lis)
    rng = np.random.default_rng(20261204)
    fd_steps = rng.integers(0, T, 32).astype(np.int32)
    fd_tris = rng.integers(0, nt, 32).astype(np.int32)
    fd = o.linear_fd(g, target, lis, fd_steps, fd_tris, 1e-3)
    s_steps = rng.integers(0, T, 4096).astype(np.int32)
    s_tris = rng.integers(0, nt, 4096).astype(np.int32)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "c4_adjoint.npz"),
                        target=target.astype(np.float64), loss=loss, grad_norm=np.linalg.norm(grad),
                        grad_maxabs=np.max(np.abs(grad)), fd_steps=fd_steps, fd_tris=fd_tris, fd=fd,
                        fd_adjoint=grad[fd_steps, fd_tris], sample_steps=s_steps, sample_tris=s_tris,
                        sample_grad=grad[s_steps, s_tris])
    print(f"loss {loss:.6e} |grad| {np.linalg.norm(grad):.6e} fd-vs-adjoint "
          f"{np.max(np.abs(fd - grad[fd_steps, fd_tris])) / np.max(np.abs(grad)):.2e} ({time.time() - t0:.0f} s)")


if __name__ == "__main__":
    main()

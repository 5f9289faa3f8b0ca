"""Writes tests/golden/c2_moving.npz: the FP64 oracle's listener signal over the whole moving
configuration (64^3; static subdivided cube of 768 triangles + icosphere L4 of 5120 triangles
translating 0.3 h at 200 Hz and rotating 2 rad/s; band-limited noise on half the vertices;
4000 steps; scenes.c2_moving), with update_mesh before every step as the demo loop does, plus
the per-step grid and face recompute counts.
Run from the repository root: python tests/golden/make_c2_moving.py"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "pppm-tdbem_amd"), os.path.join(ROOT, "oracle")]
import oracle  # noqa: E402
import scenes  # noqa: E402


def main(steps=4000):
    t0 = time.time()
    sc = scenes.c2_moving(steps=steps)
    o = oracle.Oracle(sc.res, sc.dl, sc.dt, sc.min_pos, sc.c, 0.83, oracle.FP64)
    o.set_mesh(sc.motion(0), sc.tris)
    sig = np.zeros(steps, np.float32)
    rec = np.zeros((steps, 2), np.int32)
    for i in range(steps):
        o.update_mesh(sc.motion(i))
        o.step(sc.neumann[i])
        sig[i] = o.read(*sc.listener)
        st = o.stats()
        rec[i] = (st["grid_recomputes"], st["face_recomputes"])
        if i % 500 == 499:
            print(f"step {i + 1}: {time.time() - t0:.0f} s", flush=True)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "c2_moving.npz"), listener=sig, recomputes=rec,
                        listener_cell=np.array(sc.listener, np.int32))
    print(f"{steps} steps, max |p| {np.abs(sig).max():.3e} ({time.time() - t0:.0f} s)")


if __name__ == "__main__":
    main()

"""Writes tests/golden/c1_ref32.npz: the listener signal of the static pulsating-sphere
configuration (32^3, icosphere L3, 1000 steps; scenes.c1_static_sphere) from the oracle in its
ref32 mode (the reference's FP32 geometry and complex quadrature with its double promotions,
FP16 weights, FP32 stepping) and in its FP64-weight mode, for the stated (not gated) distance
between the GPU path and the reference's own precisions.
Run from the repository root: python tests/golden/make_c1_ref32.py"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "pppm-tdbem_amd"), os.path.join(ROOT, "oracle")]
import oracle  # noqa: E402
import scenes  # noqa: E402


def run(mode, sc):
    o = oracle.Oracle(sc.res, sc.dl, sc.dt, sc.min_pos, sc.c, 0.83, mode)
    o.set_mesh(sc.verts, sc.tris)
    sig = np.zeros(sc.steps, np.float32)
    for i in range(sc.steps):
        o.step(sc.neumann[i])
        sig[i] = o.read(*sc.listener)
    return sig


def main():
    sc = scenes.c1_static_sphere(steps=1000)
    ref32, fp64 = run(oracle.REF32, sc), run(oracle.FP64, sc)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "c1_ref32.npz"), ref32=ref32, fp64=fp64)
    print("ref32 vs fp64 listener rel L2 at 300 / 1000 steps:",
          scenes.rel_l2(ref32[:300], fp64[:300]), scenes.rel_l2(ref32, fp64))


if __name__ == "__main__":
    main()

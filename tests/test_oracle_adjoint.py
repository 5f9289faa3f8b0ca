"""Pins of the oracle's FP64 step map and its adjoint (oracle/adjoint_oracle.cpp), on the CPU.

- The FP64 forward of the linear model equals the oracle's reference-precision step (FP32
  storage, pppm.cu / fdtd.cu order) to FP32 rounding: it is the same map.
- The adjoint equals central finite differences of J = sum (out - target)^2 (eps = 1e-3, FP64):
  J is quadratic in the inputs, so the central difference has no truncation error and the two
  agree to FP64 rounding. A dropped term, a wrong sign or a transposed index in the adjoint (ABC
  shell, ghost cotangents, far / Dirichlet / near transposes) shows up here.
- The dot-product identity <gout, F g> = <F^T gout, g> over random inputs and cotangents.
"""
import numpy as np
import pytest


def _scene(oracle_mod, scenes_mod, n, lv, T):
    sc = scenes_mod.c3_scene(n, lv, seed=n, steps=T)
    o = oracle_mod.Oracle(sc.res, sc.dl, sc.dt, sc.min_pos, sc.c, 0.83, oracle_mod.FP64)
    o.set_mesh(sc.verts, sc.tris)
    # one listener away from the object, one in the corrected cells next to it
    tri = o.triangles()
    near = tuple(int(x) for x in tri[0, 7:10] + np.array([1, 0, 0]))
    return sc, o, [sc.listener, near]


@pytest.mark.parametrize("n,lv", [(24, 1), (20, 2)])
def test_linear_forward_is_the_step(oracle_mod, scenes_mod, n, lv):
    T = 80
    sc, o, lis = _scene(oracle_mod, scenes_mod, n, lv, T)
    g = np.random.default_rng(1).standard_normal((T, len(sc.tris)))
    out = o.linear_forward(g, lis)
    ref = oracle_mod.Oracle(sc.res, sc.dl, sc.dt, sc.min_pos, sc.c, 0.83, oracle_mod.FP64)
    ref.set_mesh(sc.verts, sc.tris)
    r = []
    for i in range(T):
        ref.step(g[i].astype(np.float32))
        r.append([ref.read(*x) for x in lis])
    r = np.array(r)
    err = np.linalg.norm(out - r) / np.linalg.norm(r)
    assert err < 1e-5, err
    assert np.abs(r[:, 1]).max() > 0 and np.abs(r[:, 0]).max() > 0


@pytest.mark.parametrize("n,lv", [(24, 1), (20, 2)])
def test_adjoint_matches_central_differences(oracle_mod, scenes_mod, n, lv):
    T = 80
    sc, o, lis = _scene(oracle_mod, scenes_mod, n, lv, T)
    rng = np.random.default_rng(2)
    g = rng.standard_normal((T, len(sc.tris)))
    target = o.linear_forward(rng.standard_normal((T, len(sc.tris))), lis)
    grad = o.linear_adjoint(2 * (o.linear_forward(g, lis) - target), lis)
    steps = rng.integers(0, T, 24)
    tris = rng.integers(0, len(sc.tris), 24)
    fd = o.linear_fd(g, target, lis, steps, tris, 1e-3)
    scale = np.max(np.abs(grad))
    assert scale > 0
    assert np.max(np.abs(fd - grad[steps, tris])) <= 1e-8 * scale
    # the last steps' inputs cannot reach the far listener in time: their gradient there is 0,
    # but the near listener sees the last input through the near field at once
    assert np.abs(grad[-1]).max() > 0


def test_adjoint_dot_product_identity(oracle_mod, scenes_mod):
    T = 60
    sc, o, lis = _scene(oracle_mod, scenes_mod, 20, 2, T)
    rng = np.random.default_rng(3)
    for _ in range(3):
        g = rng.standard_normal((T, len(sc.tris)))
        w = rng.standard_normal((T, len(lis)))
        lhs = np.sum(w * o.linear_forward(g, lis))
        rhs = np.sum(o.linear_adjoint(w, lis) * g)
        assert abs(lhs - rhs) <= 1e-10 * (abs(lhs) + np.abs(w).sum() * 1e-12)

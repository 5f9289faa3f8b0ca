"""Pins of the oracle's modal sound source (RigidBody, src/utils/RigidBody.cu) against closed
forms: the IIR coefficients of a damped oscillator, its impulse response, the second-difference
acceleration of a sinusoid, a rigid-translation mode's surface projection, and the surface
accelerations as a matrix product."""
import math

import numpy as np
import pytest

import oracle as oracle_mod

DT = 1.0 / 44100.0


def test_undamped_coefficients_closed_form():
    freqs = np.array([50.0, 440.0, 1000.0, 3000.0, 9000.0])
    om = 2 * np.pi * freqs
    c = oracle_mod.modal_coeffs((om ** 2).astype(np.float32), DT, 0.0, 0.0).astype(np.float64)
    th = om * DT
    np.testing.assert_allclose(c[:, 0], 2 * np.cos(th), rtol=2e-6)
    np.testing.assert_array_equal(c[:, 1], -1.0)
    np.testing.assert_allclose(c[:, 2], 2 * (np.cos(th) - np.cos(2 * th)) / (3 * om ** 2), rtol=1e-4)


def test_damped_recursion_roots():
    """z^2 = c1 z + c2 has roots eps e^{+-i theta}: modulus exp(-ksi omega dt), angle omega' dt."""
    alpha, beta = 30.0, 1e-6
    freqs = np.array([200.0, 1500.0, 6000.0])
    om = 2 * np.pi * freqs
    c = oracle_mod.modal_coeffs((om ** 2).astype(np.float32), DT, alpha, beta).astype(np.float64)
    ksi = (alpha + beta * om ** 2) / (2 * om)
    eps = np.sqrt(-c[:, 1])
    np.testing.assert_allclose(eps, np.exp(-ksi * om * DT), rtol=1e-6)
    np.testing.assert_allclose(np.arccos(c[:, 0] / (2 * eps)), om * np.sqrt(1 - ksi ** 2) * DT, rtol=1e-4)


def test_impulse_response_closed_form():
    """q_n = c3 f0 eps^n sin((n+1) theta) / sin(theta) after a unit force at sample 0."""
    alpha, beta = 20.0, 2e-7
    f = 700.0
    om = 2 * np.pi * f
    c = oracle_mod.modal_coeffs(np.array([om ** 2], np.float32), DT, alpha, beta)
    eps = math.sqrt(-float(c[0, 1]))
    th = math.acos(float(c[0, 0]) / (2 * eps))
    f0 = 3.0
    S = 1500
    forces = np.zeros((S, 1), np.float32)
    forces[0, 0] = f0
    state = np.zeros((1, 3), np.float32)
    q = oracle_mod.modal_iir(c, DT, state, forces)
    n = np.arange(S)
    qn = float(c[0, 2]) * f0 * eps ** n * np.sin((n + 1) * th) / math.sin(th)
    scale = np.abs(qn).max()
    assert abs(state[0, 0] - qn[-1]) <= 1e-4 * scale
    assert abs(state[0, 1] - qn[-2]) <= 1e-4 * scale
    # the acceleration is the second difference of q over dt^2, centred on q_{n-1}
    acc = (qn[2:] - 2 * qn[1:-1] + qn[:-2]) / DT ** 2
    np.testing.assert_allclose(q[2:, 0], acc, atol=2e-4 * np.abs(acc).max())


def test_undamped_acceleration_is_minus_omega2_q():
    """For a free undamped oscillation Q_n = -(2 - 2 cos theta) / dt^2 q_{n-1} -> -omega^2 q."""
    f = 300.0
    om = 2 * np.pi * f
    c = oracle_mod.modal_coeffs(np.array([om ** 2], np.float32), DT, 0.0, 0.0)
    state = np.array([[1.0, math.cos(om * DT), math.cos(2 * om * DT)]], np.float32)  # q1, q2, q3
    forces = np.zeros((400, 1), np.float32)
    prev = []
    qq = []
    for s in range(400):
        prev.append(float(state[0, 0]))
        q = oracle_mod.modal_iir(c, DT, state, forces[s:s + 1])
        qq.append(float(q[0, 0]))
    prev = np.array(prev)
    ratio = np.array(qq)[np.abs(prev) > 0.2] / prev[np.abs(prev) > 0.2]
    th = om * DT
    np.testing.assert_allclose(ratio, -(2 - 2 * math.cos(th)) / DT ** 2, rtol=2e-3)
    np.testing.assert_allclose(ratio, -om ** 2, rtol=2e-3)


@pytest.fixture(scope="module")
def mesh():
    rng = np.random.default_rng(7)
    nv, nt, M = 60, 90, 7
    tris = np.stack([rng.choice(nv, 3, replace=False) for _ in range(nt)]).astype(np.int32)
    n = rng.standard_normal((nt, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    eig = rng.standard_normal((3 * nv, M)).astype(np.float32)
    return tris, n.astype(np.float32), eig


def test_rigid_translation_projects_to_normal_component(mesh):
    tris, n, _ = mesh
    nv = 60
    d = np.array([0.3, -1.2, 0.7], np.float32)
    eig = np.tile(d, nv)[:, None].astype(np.float32)  # every vertex moves by d
    m = oracle_mod.modal_surface_matrix(eig, tris, n)
    np.testing.assert_allclose(m[:, 0], n.astype(np.float64) @ d.astype(np.float64), atol=1e-6)


def test_surface_matrix_is_mean_vertex_displacement_dot_normal(mesh):
    tris, n, eig = mesh
    m = oracle_mod.modal_surface_matrix(eig, tris, n)
    u = eig.astype(np.float64).reshape(-1, 3, eig.shape[1])       # (nv, 3, M)
    mean = u[tris].mean(axis=1)                                   # (nt, 3, M)
    ref = np.einsum("tdm,td->tm", mean, n.astype(np.float64))
    np.testing.assert_allclose(m, ref, atol=2e-6 * np.abs(ref).max())


def test_accelerations_are_the_matrix_product(mesh):
    rng = np.random.default_rng(9)
    tris, n, eig = mesh
    m = oracle_mod.modal_surface_matrix(eig, tris, n)
    q = rng.standard_normal((33, eig.shape[1])).astype(np.float32)
    acc = oracle_mod.modal_accs(m, q)
    ref = q.astype(np.float64) @ m.astype(np.float64).T
    np.testing.assert_allclose(acc, ref, atol=1e-5 * np.abs(ref).max())


def test_force_of_an_impulse_at_its_peak(mesh):
    """At start + tau/2 the contact pulse sin(pi (t - start) / tau) is 1: f_i = E(v, i) . I."""
    _, _, eig = mesh
    t0 = np.float32(0.01)
    tau = np.float32(0.0005)
    vec = np.array([[0.2, -0.5, 1.5]], np.float32)
    f = oracle_mod.modal_forces(eig, [t0], [13], vec, [1], float(t0 + tau / 2))
    ref = eig[39:42].astype(np.float64).T @ vec[0].astype(np.float64)
    np.testing.assert_allclose(f, ref, rtol=1e-5, atol=1e-5)
    off = oracle_mod.modal_forces(eig, [t0], [13], vec, [0], float(t0 + tau / 2))
    np.testing.assert_array_equal(off, 0.0)


def test_accumulation_is_fused_per_term():
    """Q = (-(1 + 2^-11), 1 + 2^-12), M = (1, 1 + 2^-12): the exact second product is
    1 + 2^-11 + 2^-24, so a fused multiply-add leaves 2^-24 while a rounded product (a tie, to
    even) cancels to 0 — the device compiler's contraction of `acc += M * Q` is the former."""
    m = np.array([[1.0, 1.0 + 2.0 ** -12]], np.float32)
    q = np.array([[-(1.0 + 2.0 ** -11), 1.0 + 2.0 ** -12]], np.float32)
    acc = oracle_mod.modal_accs(m, q)
    assert acc[0, 0] == np.float32(2.0 ** -24)
    # in mode order: the same two terms the other way round start from the inexact product
    acc_rev = oracle_mod.modal_accs(m[:, ::-1].copy(), q[:, ::-1].copy())
    assert acc_rev[0, 0] == np.float32(0.0)

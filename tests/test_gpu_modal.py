"""GPU modal sound source (pppm_modal_* through the C ABI) against the oracle: IIR coefficients,
surface modal matrix, oscillator state and surface accelerations bit for bit, over two calls in
a row (host forces, then device forces), so the oscillator state carries over."""
import numpy as np
import pytest

import oracle as oracle_mod

pytestmark = pytest.mark.gpu


def _problem(seed=3, nv=400, nt=700, M=48):
    rng = np.random.default_rng(seed)
    tris = np.stack([rng.choice(nv, 3, replace=False) for _ in range(nt)]).astype(np.int32)
    n = rng.standard_normal((nt, 3))
    n = (n / np.linalg.norm(n, axis=1, keepdims=True)).astype(np.float32)
    eig = (rng.standard_normal((3 * nv, M)) * 0.1).astype(np.float32)
    freqs = np.sort(rng.uniform(150.0, 9000.0, M))
    eigvals = ((2 * np.pi * freqs) ** 2).astype(np.float32)
    return tris, n, eig, eigvals, rng


def test_modal_source_bitexact():
    import torch
    import pppm_amd
    tris, n, eig, eigvals, rng = _problem()
    alpha, beta, rate = 30.0, 1e-6, 44100.0
    ms = pppm_amd.ModalSource(eig, eigvals, tris, n, alpha, beta, rate)
    dt = np.float32(1.0) / np.float32(rate)

    c_ref = oracle_mod.modal_coeffs(eigvals, float(dt), alpha, beta)
    assert np.array_equal(ms.coeffs().view(np.uint32), c_ref.view(np.uint32))
    m_ref = oracle_mod.modal_surface_matrix(eig, tris, n)
    assert np.array_equal(ms.surface_matrix().view(np.uint32), m_ref.view(np.uint32))

    state = np.zeros((eig.shape[1], 3), np.float32)
    for S, on_device in ((777, False), (64, True)):
        forces = (rng.standard_normal((S, eig.shape[1])) * 1e3).astype(np.float32)
        accs = torch.zeros((S, tris.shape[0]), dtype=torch.float32, device="cuda")
        ms.run(torch.from_numpy(forces).cuda() if on_device else forces, accs)
        torch.cuda.synchronize()
        q = oracle_mod.modal_iir(c_ref, float(dt), state, forces)
        a_ref = oracle_mod.modal_accs(m_ref, q)
        got = accs.cpu().numpy()
        diff = int(np.count_nonzero(got.view(np.uint32) != a_ref.view(np.uint32)))
        print(f"S={S} on_device={on_device}: differing accelerations {diff} of {got.size}", flush=True)
        assert diff == 0
        assert np.array_equal(ms.state().view(np.uint32), state.view(np.uint32))


def test_modal_state_round_trip():
    import torch
    import pppm_amd
    tris, n, eig, eigvals, rng = _problem(seed=5, nv=50, nt=80, M=9)
    ms = pppm_amd.ModalSource(eig, eigvals, tris, n, 10.0, 1e-7, 44100.0)
    forces = rng.standard_normal((100, 9)).astype(np.float32)
    accs = torch.zeros((100, 80), dtype=torch.float32, device="cuda")
    ms.run(forces, accs)
    saved = ms.state()
    a1 = torch.zeros_like(accs)
    ms.run(forces, a1)
    ms.set_state(saved)
    a2 = torch.zeros_like(accs)
    ms.run(forces, a2)
    torch.cuda.synchronize()
    assert torch.equal(a1, a2)

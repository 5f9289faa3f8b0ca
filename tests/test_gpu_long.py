"""Whole-length runs against committed oracle fixtures (tests/golden/make_c2_moving.py,
tests/golden/make_c1_ref32.py):
- the moving configuration (64^3, cube + icosphere L4, update_mesh before every step) for all
  4000 steps against the FP64 oracle's listener: the 1e-5 relative L2 gate;
- the static pulsating sphere (32^3, 1000 steps) against the oracle in ref32 mode (the
  reference's own precisions): the distance is stated, not gated (the FP16 codes of ref32 and
  FP64-built weights differ by one ulp in a few records and the scheme amplifies it)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
LISTEN_TOL = 1e-5


def test_c2_moving_4000_steps_against_the_oracle(amd, scenes_mod):
    import torch
    ref = np.load(os.path.join(HERE, "golden", "c2_moving.npz"))
    steps = len(ref["listener"])
    sc = scenes_mod.c2_moving(steps=steps)
    assert tuple(ref["listener_cell"]) == tuple(sc.listener)
    print(f"c2 moving, {steps} steps vs the FP64 oracle", flush=True)
    verts = torch.from_numpy(np.stack([sc.motion(i) for i in range(steps)]).astype(np.float32)).cuda()
    nm = torch.from_numpy(sc.neumann).cuda()
    g = amd.PPPMSolver(sc.res, sc.dl, sc.dt, sc.min_pos, 0, 0.83, sc.c)
    g.set_mesh(sc.motion(0), sc.tris)
    out = torch.zeros(steps, 1, device="cuda")
    g.run_moving(verts, nm, steps, [sc.listener], out)
    sig = out[:, 0].cpu().numpy()
    want = ref["listener"]
    err = scenes_mod.rel_l2(sig, want)
    same = int(np.sum(sig.view(np.uint32) == want.view(np.uint32)))
    first = int(np.argmax(sig.view(np.uint32) != want.view(np.uint32))) if same < steps else steps
    print(f"  listener rel L2 {err:.2e}; {same} of {steps} samples bit-identical (first difference at {first})")
    assert np.abs(want).max() > 0 and err < LISTEN_TOL


def test_c1_distance_to_the_reference_precisions(amd, scenes_mod):
    import torch
    ref = np.load(os.path.join(HERE, "golden", "c1_ref32.npz"))
    sc = scenes_mod.c1_static_sphere(steps=len(ref["ref32"]))
    g = amd.PPPMSolver(sc.res, sc.dl, sc.dt, sc.min_pos, 0, 0.83, sc.c)
    g.set_mesh(sc.verts, sc.tris)
    out = torch.zeros(sc.steps, 1, device="cuda")
    g.run(torch.from_numpy(sc.neumann).cuda(), sc.steps, [sc.listener], out)
    sig = out[:, 0].cpu().numpy()
    e64 = scenes_mod.rel_l2(sig, ref["fp64"])
    e32 = [scenes_mod.rel_l2(sig[:n], ref["ref32"][:n]) for n in (300, 1000)]
    print(f"c1 listener: vs FP64 oracle {e64:.2e} (gated 1e-5); vs ref32 oracle (stated) "
          f"{e32[0]:.2e} after 300 steps, {e32[1]:.2e} after 1000")
    assert e64 < LISTEN_TOL
    assert np.isfinite(e32).all()

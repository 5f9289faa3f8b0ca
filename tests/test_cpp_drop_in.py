"""The C++ host interface (include/pppm_solver.hpp) used the way the reference's drivers use
their PPPMSolver, through a small compiled driver (tests/cpp/drop_in.cpp)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
EXE = os.path.join(CPP, "_build", "drop_in")


def _build():
    subprocess.run(["make", "-C", CPP, "-s"], check=True)


def test_cpp_interface_raises_on_bad_arguments():
    _build()
    r = subprocess.run([EXE, "args"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.startswith("ok"), (r.stdout, r.stderr)


@pytest.mark.gpu
def test_cpp_driver_matches_oracle(tmp_path, oracle_mod, scenes_mod):
    """Static sphere on 32^3, 120 steps through the C++ interface: listener bit-exact with
    the FP64 oracle (the GPU path reproduces its operation order)."""
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (make -C tests/cpp)")
    steps = 120
    sc = scenes_mod.c1_static_sphere(steps=steps)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([sc.res, sc.verts.shape[0], sc.tris.shape[0], steps, *sc.listener], np.int32).tofile(f)
        np.array([sc.dl, sc.dt, sc.c, *sc.min_pos], np.float32).tofile(f)
        sc.verts.astype(np.float32).tofile(f)
        sc.tris.astype(np.int32).tofile(f)
        sc.neumann[:steps].astype(np.float32).tofile(f)
    print(f"cpp drop-in: C1 {steps} steps", flush=True)
    r = subprocess.run([EXE, "run", str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.stdout, r.stderr)
    got = np.fromfile(out, np.float32)
    o = oracle_mod.Oracle(sc.res, sc.dl, sc.dt, sc.min_pos, sc.c, 0.83, oracle_mod.FP64)
    o.set_mesh(sc.verts, sc.tris)
    ref = []
    for i in range(steps):
        o.step(sc.neumann[i])
        ref.append(o.read(*sc.listener))
    assert np.array_equal(got, np.array(ref, np.float32))

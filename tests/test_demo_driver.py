"""The demo driver (pppm-tdbem_amd/driver): the reference's scene config, Audio/Manual objects,
per-sample loop with chunked restart and cross-fade, and result.txt (experiments/demo/main.cu).

CPU: the derived grid/timing parameters and the recorded solver-call sequence are checked
against independent computations (scipy rotations, the acceleration sample index, the
restart schedule). GPU: the driver's result.txt equals a replay of its own call sequence on
the oracle."""
import math
import os
import struct
import subprocess

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRV = os.path.join(ROOT, "pppm-tdbem_amd", "driver")
EXE = os.path.join(DRV, "_build", "pppm_demo")


def _build():
    subprocess.run(["make", "-C", DRV, "-s"], check=True)


def write_obj(path, verts, tris):
    with open(path, "w") as f:
        for v in verts:
            f.write(f"v {v[0]:.9g} {v[1]:.9g} {v[2]:.9g}\n")
        for t in tris:
            f.write(f"f {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")


def make_scene(tmp, scenes_mod, res=24, max_time=0.004, chunk=40, back=10, frames=30, seed=5):
    """A moving icosphere (Audio: motion, accelerations, selected vertices) beside a static
    subdivided cube (Manual), in the reference's directory and config layout."""
    rng = np.random.default_rng(seed)
    inp, out = os.path.join(tmp, "input"), os.path.join(tmp, "output")
    os.makedirs(os.path.join(inp, "ball"))
    os.makedirs(os.path.join(inp, "box"))
    v, t = scenes_mod.icosphere(2, 0.02, (0.0, 0.0, 0.0))
    write_obj(os.path.join(inp, "ball", "mesh.obj"), v, t)
    cv, ct = scenes_mod.subdivided_cube(1, 0.03, (0.06, 0.0, 0.0))
    write_obj(os.path.join(inp, "box", "mesh.obj"), cv, ct)
    times = np.arange(frames) * (max_time / (frames - 4))
    axis = rng.standard_normal(3)
    axis /= np.linalg.norm(axis)
    with open(os.path.join(inp, "ball", "motion.txt"), "w") as f:
        for k, tt in enumerate(times):
            q = Rotation.from_rotvec(axis * 40.0 * tt).as_quat()  # x, y, z, w
            tr = (-0.03 + 0.002 * math.sin(2 * math.pi * 300 * tt), 0.001 * k / frames, 0.0)
            f.write(f"{tt:.9g} {tr[0]:.9g} {tr[1]:.9g} {tr[2]:.9g} {q[0]:.9g} {q[1]:.9g} {q[2]:.9g} {q[3]:.9g}\n")
    n_acc = int(max_time * 44100) + 50
    accs = np.sin(2 * np.pi * 2000 * np.arange(n_acc) / 44100) * 50.0
    with open(os.path.join(inp, "ball", "accs.txt"), "w") as f:
        f.write("\n".join(f"{a:.7g}" for a in accs) + "\n")
    sel = rng.choice(len(v), 40, replace=False)
    with open(os.path.join(inp, "ball", "selected_vertices.txt"), "w") as f:
        f.write("\n".join(str(s) for s in sel) + "\n")
    cfg = os.path.join(tmp, "scene.cfg")
    with open(cfg, "w") as f:
        f.write("{\n")
        f.write(f'    inputDir: "{inp}",\n    outputDir: "{out}",\n')
        f.write(f"    res: {res},\n    gridLengthFactor: 2,\n    reflectCoeff: 0.83,\n")
        f.write(f"    maxTime: {max_time},\n    chunkSize: {chunk},\n    backStepSize: {back},\n")
        f.write("    checkCoordX: 0.875,\n    checkCoordY: 0.875,\n    checkCoordZ: 0.875,\n")
        f.write("    sceneInfo: [\n")
        f.write('        {\n            name: "ball",\n            type: "Audio"\n        },\n')
        f.write('        {\n            name: "box",\n            type: "Manual"\n        }\n')
        f.write("    ]\n}\n")
    return dict(cfg=cfg, out=out, ball=(v, t), box=(cv, ct), times=times, axis=axis, accs=accs, sel=sel,
                frames=frames, max_time=max_time)


def read_calls(path):
    data = open(path, "rb").read()
    o, calls = 0, []
    nv = nt = 0
    while o < len(data):
        (op,) = struct.unpack_from("i", data, o)
        o += 4
        if op == 0:
            nv, nt = struct.unpack_from("ii", data, o)
            o += 8
            v = np.frombuffer(data, np.float32, 3 * nv, o).reshape(nv, 3)
            o += 12 * nv
            t = np.frombuffer(data, np.int32, 3 * nt, o).reshape(nt, 3)
            o += 12 * nt
            calls.append(("set_mesh", v, t))
        elif op == 1:
            v = np.frombuffer(data, np.float32, 3 * nv, o).reshape(nv, 3)
            o += 12 * nv
            calls.append(("update_mesh", v))
        elif op == 2:
            n = np.frombuffer(data, np.float32, nt, o)
            o += 4 * nt
            calls.append(("step", n))
        elif op == 3:
            idx, w = struct.unpack_from("if", data, o)
            o += 8
            calls.append(("read", idx, w))
        elif op == 4:
            (t,) = struct.unpack_from("i", data, o)
            o += 4
            calls.append(("reset", t))
        else:
            raise AssertionError(f"bad op {op}")
    return calls


def setup_of(cfg):
    out = subprocess.run([EXE, "--setup", cfg], capture_output=True, text=True, check=True).stdout
    return {ln.split()[0]: ln.split()[1:] for ln in out.strip().splitlines()}


def test_setup_follows_the_scene(tmp_path, scenes_mod):
    _build()
    sc = make_scene(str(tmp_path), scenes_mod)
    s = setup_of(sc["cfg"])
    v, _ = sc["ball"]
    rots = Rotation.from_quat([Rotation.from_rotvec(sc["axis"] * 40.0 * t).as_quat() for t in sc["times"]])
    pts = [np.asarray(sc["box"][0], np.float64)]
    for k, t in enumerate(sc["times"]):
        tr = np.array([-0.03 + 0.002 * math.sin(2 * math.pi * 300 * t), 0.001 * k / sc["frames"], 0.0])
        pts.append(rots[k].apply(v.astype(np.float64)) + tr)
    pts = np.concatenate(pts)
    lo, hi = pts.min(0), pts.max(0)
    length = (hi - lo).max() * 2
    h = length / 24
    assert float(s["grid_size"][0]) == pytest.approx(h, rel=1e-5)
    assert np.allclose([float(x) for x in s["min_pos"]], (lo + hi) / 2 - length / 2, atol=1e-6)
    rate = int(1.01 / (h / math.sqrt(3) / 340.3))
    assert abs(int(s["frame_rate"][0]) - rate) <= 1
    assert int(s["frame_num"][0]) in (int(sc["max_time"] * rate) - 1, int(sc["max_time"] * rate))
    assert [int(x) for x in s["check"]] == [21, 21, 21]
    assert int(s["triangles"][0]) == len(sc["ball"][1]) + len(sc["box"][1])


def test_recorded_calls_follow_motion_accelerations_and_restarts(tmp_path, scenes_mod):
    _build()
    sc = make_scene(str(tmp_path), scenes_mod)
    s = setup_of(sc["cfg"])
    rec = os.path.join(str(tmp_path), "calls.bin")
    subprocess.run([EXE, "--calls", sc["cfg"], rec], check=True)
    calls = read_calls(rec)
    dt = np.float32(1.0) / np.float32(int(s["frame_rate"][0]))
    frame_num = int(s["frame_num"][0])
    v, t = sc["ball"]
    nb = len(v)
    chunk, back = 40, 10

    # the restart schedule of main.cu:209-263: plain steps, and at i % chunk == chunk - back a
    # silent fade-out pass, a reset to the saved time index, and a fade-in pass
    plan, i = [], 0
    while i <= frame_num:
        if i % chunk == chunk - back:
            end = min(i + back, frame_num + 1)
            plan += [("out", j) for j in range(i, end)] + [("reset", i)] + [("in", j) for j in range(i, end)]
            i = end
        else:
            plan.append(("plain", i))
            i += 1
    reads = [c for c in calls if c[0] in ("read", "reset")]
    assert len(reads) == len(plan)
    for (kind, j), c in zip(plan, reads):
        if kind == "reset":
            assert c[0] == "reset" and c[1] == j - 1      # fdtd.t after j steps is j - 1
            continue
        assert c[0] == "read" and c[1] == j + 1
        if kind == "plain":
            assert c[2] == 1.0
        elif kind == "in":
            assert c[2] < 0
        else:
            # fade-out weight 1 - (j - i) / (end - i)
            i0 = j - (j % chunk - (chunk - back))
            end = min(i0 + back, frame_num + 1)
            assert c[2] == np.float32(1) - np.float32(j - i0) / np.float32(end - i0)

    # accelerations: sample K = max{k : k / 44100 <= time} (float32) on the triangles closest to
    # the selected vertices (and vertex 0); silent in the fade-out pass
    accs = np.array([np.float32(x) for x in open(os.path.join(os.path.dirname(sc["cfg"]), "input", "ball",
                                                                "accs.txt")).read().split()])
    st = np.float32(1) / np.float32(44100)
    sel = set(int(np.flatnonzero(np.any(t == vid, axis=1))[0]) for vid in list(sc["sel"]) + [0])
    on = np.zeros(len(t) + len(sc["box"][1]), bool)
    on[sorted(sel)] = True
    steps = [c for c in calls if c[0] == "step"]
    assert len(steps) == sum(1 for k, _ in plan if k != "reset")
    curr, saved, n = np.float32(0), np.float32(0), 0
    for kind, j in plan:
        if kind == "reset":
            curr = saved
            continue
        if kind == "out" and j % chunk == chunk - back:
            saved = curr
        curr = np.float32(curr + dt)
        x = steps[n][1]
        n += 1
        if kind == "out":
            assert not np.any(x)
            continue
        K = int(np.floor(np.float64(curr) * 44100))
        while np.float32(st * np.float32(K + 1)) <= curr:
            K += 1
        while K > 0 and np.float32(st * np.float32(K)) > curr:
            K -= 1
        assert np.all(x[~on] == 0)
        assert np.all(x[on] == accs[K]), (kind, j, K)

    # meshes: the ball is its rest shape under the rotation and translation of the frame in
    # effect (last frame time <= current time), the box never moves
    times = np.array([np.float32(ln.split()[0]) for ln in open(os.path.join(os.path.dirname(sc["cfg"]), "input",
                                                                             "ball", "motion.txt"))])
    mot = np.loadtxt(os.path.join(os.path.dirname(sc["cfg"]), "input", "ball", "motion.txt"))
    for c in calls:
        if c[0] in ("set_mesh", "update_mesh"):
            assert np.array_equal(c[1][nb:], np.asarray(sc["box"][0], np.float32))
            ball = c[1][:nb].astype(np.float64)
            errs = [np.max(np.abs(Rotation.from_quat(m[4:8]).apply(v.astype(np.float64)) + m[1:4] - ball))
                    for m in mot]
            assert min(errs) < 1e-6
    assert sum(1 for c in calls if c[0] == "update_mesh") > 0


@pytest.mark.gpu
def test_driver_result_matches_oracle_replay(tmp_path, scenes_mod, oracle_mod):
    """result.txt of a GPU run equals the oracle's replay of the driver's own call record."""
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (make -C pppm-tdbem_amd/driver)")
    sc = make_scene(str(tmp_path), scenes_mod)
    s = setup_of(sc["cfg"])
    rec = os.path.join(str(tmp_path), "calls.bin")
    subprocess.run([EXE, "--calls", sc["cfg"], rec], check=True)
    print("demo driver: GPU run", flush=True)
    r = subprocess.run([EXE, sc["cfg"]], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout, r.stderr)
    got = np.loadtxt(os.path.join(sc["out"], "result.txt"), dtype=np.float64)
    res = int(s["res"][0])
    h = float(s["grid_size"][0])
    dt = np.float32(1.0) / np.float32(int(s["frame_rate"][0]))
    mn = [float(x) for x in s["min_pos"]]
    check = [int(x) for x in s["check"]]
    o = None
    out = np.zeros(len(got), np.float32)
    out[0] = np.float32(int(s["frame_rate"][0]))
    for c in read_calls(rec):
        if c[0] == "set_mesh":
            o = oracle_mod.Oracle(res, h, float(dt), mn, 340.3, 0.83, oracle_mod.FP64)
            o.set_mesh(c[1], c[2])
        elif c[0] == "update_mesh":
            o.update_mesh(c[1])
        elif c[0] == "step":
            o.step(c[1])
        elif c[0] == "read":
            p = np.float32(o.read(*check))
            if c[2] < 0:
                out[c[1]] = np.float32(out[c[1]] + p)
            else:
                out[c[1]] = np.float32(p * np.float32(c[2]))
        elif c[0] == "reset":
            o.reset()
            o.set_time_idx(c[1])
    ref = np.array([float(f"{x:e}") for x in out])
    assert np.array_equal(got, ref)


# ------------------------------------------------------------------ modal objects
def _kuhn_cube(n, a):
    """[-a, a]^3 as n^3 cubes of 6 tetrahedra each: vertices (x-major) and tetrahedra."""
    xs = np.linspace(-a, a, n + 1)
    idx = lambda i, j, k: (i * (n + 1) + j) * (n + 1) + k
    verts = np.array([[xs[i], xs[j], xs[k]] for i in range(n + 1) for j in range(n + 1) for k in range(n + 1)],
                     np.float32)
    tets = []
    for i in range(n):
        for j in range(n):
            for k in range(n):
                for perm in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
                    c = [0, 0, 0]
                    path = [idx(i, j, k)]
                    for ax in perm:
                        c[ax] = 1
                        path.append(idx(i + c[0], j + c[1], k + c[2]))
                    tets.append(path)
    return verts, np.array(tets, np.int32)


def make_modal_scene(tmp, scenes_mod, res=24, max_time=0.003, chunk=40, back=10):
    """A modal object (a cube of 48 tetrahedra with synthetic eigenpairs, two contact impulses
    and a slow translation) beside a static subdivided cube (Manual)."""
    rng = np.random.default_rng(8)
    inp, out = os.path.join(tmp, "input"), os.path.join(tmp, "output")
    os.makedirs(os.path.join(inp, "bell"))
    os.makedirs(os.path.join(inp, "box"))
    v, tets = _kuhn_cube(2, 0.02)
    d = os.path.join(inp, "bell")
    np.savetxt(os.path.join(d, "vertices.txt"), v, fmt="%.9g")
    np.savetxt(os.path.join(d, "tets.txt"), tets, fmt="%d")
    freqs = np.array([400.0, 900.0, 1500.0, 2600.0, 4000.0, 7000.0, 16000.0])  # the last is cut off
    eigvals = ((2 * np.pi * freqs) ** 2).astype(np.float32)
    np.savetxt(os.path.join(d, "eigenvalues.txt"), eigvals, fmt="%.9g")
    eig = (rng.standard_normal((3 * len(v), len(freqs))) * 0.1).astype(np.float32)
    np.savetxt(os.path.join(d, "eigenvectors.txt"), eig, fmt="%.9g")
    imps = [(0.0015, 26, (0.0, 0.0, 1.0), 2.0), (0.00162, 8, (1.0, 0.0, 0.0), 1.0)]
    with open(os.path.join(d, "contact.txt"), "w") as f:
        for t, vid, n, m in imps:
            f.write(f"{t} {vid} {n[0]} {n[1]} {n[2]} {m}\n")
    with open(os.path.join(d, "motion.txt"), "w") as f:
        for k in range(8):
            f.write(f"{0.001 * k:.9g} {0.0005 * k:.9g} 0 0 0 0 0 1\n")
    cv, ct = scenes_mod.subdivided_cube(1, 0.03, (0.07, 0.0, 0.0))
    write_obj(os.path.join(inp, "box", "mesh.obj"), cv, ct)
    cfg = os.path.join(tmp, "scene.cfg")
    with open(cfg, "w") as f:
        f.write("{\n")
        f.write(f'    inputDir: "{inp}",\n    outputDir: "{out}",\n')
        f.write(f"    res: {res},\n    gridLengthFactor: 2,\n    reflectCoeff: 0.83,\n")
        f.write(f"    maxTime: {max_time},\n    chunkSize: {chunk},\n    backStepSize: {back},\n")
        f.write("    checkCoordX: 0.875,\n    checkCoordY: 0.875,\n    checkCoordZ: 0.875,\n")
        f.write("    sceneInfo: [\n")
        f.write('        {\n            name: "bell",\n            type: "Modal"\n        },\n')
        f.write('        {\n            name: "box",\n            type: "Manual"\n        }\n')
        f.write("    ]\n}\n")
    return dict(cfg=cfg, out=out, verts=v, eig=eig, eigvals=eigvals, imps=imps, chunk=chunk, back=back,
                max_time=max_time)


class _ModalEmu:
    """The driver's modal object (RigidBody audio_step / CalculateIIR_) on the oracle: one
    audio step per sample at timestep 1 / frame_rate, contact impulses as modal forces."""

    def __init__(self, oracle_mod, sc, tris, normals, rate):
        f32 = np.float32
        self.o = oracle_mod
        self.ts = f32(1.0) / f32(rate)
        keep = int(np.sum(np.sqrt(sc["eigvals"].astype(np.float32)) / (2 * np.pi) < 15000.0))
        self.eig = np.ascontiguousarray(sc["eig"][:, :keep])
        self.c = oracle_mod.modal_coeffs(sc["eigvals"][:keep], float(self.ts), 30.0, 1e-6)
        self.m = oracle_mod.modal_surface_matrix(self.eig, tris, normals)
        self.imp_t = np.array([f32(t) for t, _, _, _ in sc["imps"]], np.float32)
        self.imp_v = np.array([int(f32(vid) + f32(0.5)) for _, vid, _, _ in sc["imps"]], np.int32)
        self.imp_vec = np.array([np.array(n, np.float32) * f32(m) for _, _, n, m in sc["imps"]], np.float32)
        self.time, self.imp, self.active = f32(0), 0, []
        self.state = np.zeros((keep, 3), np.float32)
        self.accs = np.zeros(len(tris), np.float32)

    def step(self):
        f32 = np.float32
        self.time = f32(self.time + self.ts)
        while self.imp < len(self.imp_t) and self.time >= self.imp_t[self.imp]:
            self.active.append(self.imp)
            self.imp += 1
        while self.active and f32(self.time - self.imp_t[self.active[0]]) > f32(0.0005):
            self.active.pop(0)
        forces = np.zeros(self.eig.shape[1], np.float32)
        for j in self.active:  # one impulse at a time, in list order (duplicates included)
            forces = (forces + self.o.modal_forces(self.eig, [self.imp_t[j]], [self.imp_v[j]],
                                                   self.imp_vec[j:j + 1], [1], float(self.time))).astype(np.float32)
        q = self.o.modal_iir(self.c, float(self.ts), self.state, forces[None, :])
        self.accs = self.o.modal_accs(self.m, q)[0]

    def save(self):
        return (self.time, self.imp, self.state.copy(), self.accs.copy())

    def load(self, s):
        self.time, self.imp, st, ac = s
        self.state[:] = st
        self.accs = ac.copy()


def _normals(v, tris):
    f32 = np.float32
    e1 = v[tris[:, 1]] - v[tris[:, 0]]
    e2 = v[tris[:, 2]] - v[tris[:, 0]]
    c = np.stack([e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1], e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2],
                  e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]], 1).astype(np.float32)
    dd = (c[:, 0] * c[:, 0] + c[:, 1] * c[:, 1] + c[:, 2] * c[:, 2]).astype(np.float32)
    inv = (f32(1.0) / np.sqrt(dd)).astype(np.float32)
    return (c * inv[:, None]).astype(np.float32)


@pytest.mark.gpu
def test_driver_modal_object(tmp_path, scenes_mod, oracle_mod):
    """A Modal object in the demo driver: its surface is the tetrahedral mesh's closed, outward
    boundary; the samples before the first contact are silent; every step's accelerations on
    its triangles equal the oracle's modal restatement (forces from the contact impulses, the
    oscillators, the surface sum) bit for bit, through the chunked restart; and result.txt equals
    the oracle's replay of the recorded calls."""
    import torch  # noqa: F401  (initialises the runtime the driver shares)
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (make -C pppm-tdbem_amd/driver)")
    sc = make_modal_scene(str(tmp_path), scenes_mod)
    rec = os.path.join(str(tmp_path), "calls.bin")
    print("demo driver: modal object, GPU run with trace", flush=True)
    r = subprocess.run([EXE, "--trace", sc["cfg"], rec], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout, r.stderr)
    got = np.loadtxt(os.path.join(sc["out"], "result.txt"), dtype=np.float64)
    rate = int(got[0])
    calls = read_calls(rec)
    # surface: 6 faces x 4 squares x 2 triangles, each edge shared by two triangles, outward
    v0, t0 = calls[0][1], calls[0][2]
    tris = t0[:48]
    assert len(t0) > 48 and len(v0) > len(sc["verts"])
    edges = {}
    for t in tris:
        for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0])):
            edges[(min(a, b), max(a, b))] = edges.get((min(a, b), max(a, b)), 0) + 1
    assert len(edges) == 72 and set(edges.values()) == {2}
    nrm = _normals(sc["verts"], tris)
    cen = sc["verts"][tris].mean(axis=1)
    assert np.all(np.sum(nrm * cen, axis=1) > 0)
    # the oscillators, replayed through the driver's loop (main.cu:166-272 with the restarts)
    emu = _ModalEmu(oracle_mod, sc, tris, nrm, rate)
    moves = 0
    while f32_lt(emu.time + emu.ts, emu.imp_t[0]):
        emu.step()
        moves += 1
    mute = int(np.float32(emu.time) * np.float32(rate))
    assert np.all(got[1:mute + 1] == 0.0)
    steps = [c for c in calls if c[0] == "step"]
    expect, k = [], 0
    s_chunk, s_back = sc["chunk"], sc["back"]
    n_frames = len(got) - 2 - mute
    progress = 0
    while progress <= n_frames:
        i = progress
        if i % s_chunk == s_chunk - s_back:
            end = min(i + s_back, n_frames + 1)
            saved = emu.save()
            for _ in range(i, end):
                emu.step()
                expect.append(None)  # silent input
            emu.load(saved)
            for _ in range(i, end):
                emu.step()
                expect.append(emu.accs.copy())
                progress += 1
            continue
        emu.step()
        expect.append(emu.accs.copy())
        progress += 1
    assert len(expect) == len(steps)
    for e, st in zip(expect, steps):
        a = st[1][:48]
        if e is None:
            assert not np.any(a)
        else:
            assert np.array_equal(a.view(np.uint32), e.view(np.uint32))
    assert any(np.any(e) for e in expect if e is not None)
    # the solver side: the recorded calls replayed on the oracle, on the driver's grid
    st = setup_of(sc["cfg"])
    assert int(st["mute"][0]) == mute and int(st["frame_rate"][0]) == rate
    res = int(st["res"][0])
    gs = float(st["grid_size"][0])
    mn = [float(x) for x in st["min_pos"]]
    check = [int(x) for x in st["check"]]
    sp = np.float32(1.0) / np.float32(rate)
    o, out = None, np.zeros(len(got), np.float32)
    out[0] = np.float32(rate)
    for c in calls:
        if c[0] == "set_mesh":
            o = oracle_mod.Oracle(res, gs, float(sp), mn, 340.3, 0.83, oracle_mod.FP64)
            o.set_mesh(c[1], c[2])
        elif c[0] == "update_mesh":
            o.update_mesh(c[1])
        elif c[0] == "step":
            o.step(c[1])
        elif c[0] == "read":
            pv = np.float32(o.read(*check))
            out[c[1]] = np.float32(out[c[1]] + pv) if c[2] < 0 else np.float32(pv * np.float32(c[2]))
        elif c[0] == "reset":
            o.reset()
            o.set_time_idx(c[1])
    ref = np.array([float(f"{x:e}") for x in out])
    assert np.array_equal(got, ref)


def f32_lt(a, b):
    return bool(np.float32(a) < np.float32(b))

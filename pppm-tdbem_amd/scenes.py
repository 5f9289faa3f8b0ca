"""Procedural meshes, workloads and excitations for the PPPM hot path.

The reference loads OBJ assets (assets/sphere4.obj = icosphere level 3, assets/sphere.obj =
level 4, assets/cube.obj); this module generates the same shapes procedurally so that no
asset is needed. Workloads:

  c1_static_sphere  : icosphere L3 (642 v / 1280 t), radius 0.05 m, centred in a 32^3 grid
                      of length 0.2 m (gridLengthFactor 2 on the sphere diameter), c = 343,
                      dt = h / (sqrt(3) c), Gaussian-modulated 1 kHz normal acceleration,
                      listener at relative (0.875, 0.875, 0.875) (experiments/demo/main.cu:148-149).
  c2_moving         : 64^3 grid, static cube (12 -> 768 tris) + icosphere L4 (5120 tris)
                      translating sinusoidally (0.3 h, 200 Hz) and rotating at 2 rad/s about a
                      seeded axis; band-limited (0-4 kHz) unit-RMS white noise on a seeded 50 %
                      vertex subset, triangle value = mean of its three vertices.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

SEED = 20261202


# ---------------------------------------------------------------------- meshes
def icosphere(level: int, radius: float = 1.0, center=(0.0, 0.0, 0.0)):
    """Icosahedron subdivided `level` times, vertices projected to the sphere; outward CCW."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
         (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts) * radius + np.asarray(center, dtype=np.float64)
    return V.astype(np.float32), np.array(faces, dtype=np.int32)


def subdivided_cube(levels: int, side: float = 1.0, center=(0.0, 0.0, 0.0)):
    """Unit cube (12 outward triangles) with `levels` rounds of 4-way midpoint subdivision."""
    c = np.array([[x, y, z] for x in (-0.5, 0.5) for y in (-0.5, 0.5) for z in (-0.5, 0.5)], dtype=np.float64)
    # vertex index = 4*x + 2*y + z with x,y,z in {0,1}
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    faces = []
    for a, b, cc, d in quads:
        faces += [(a, b, cc), (a, cc, d)]
    verts = [p for p in c]
    # make every face outward
    fixed = []
    for f in faces:
        p0, p1, p2 = (verts[i] for i in f)
        n = np.cross(p1 - p0, p2 - p0)
        if np.dot(n, (p0 + p1 + p2) / 3) < 0:
            f = (f[0], f[2], f[1])
        fixed.append(f)
    faces = fixed
    for _ in range(levels):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                verts.append((verts[a] + verts[b]) / 2)
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, cc in faces:
            ab, bc, ca = mid(a, b), mid(b, cc), mid(cc, a)
            nf += [(a, ab, ca), (b, bc, ab), (cc, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts) * side + np.asarray(center, dtype=np.float64)
    return V.astype(np.float32), np.array(faces, dtype=np.int32)


# ---------------------------------------------------------------------- signals
def gaussian_sine(t, f0=1000.0, t0=4e-3, sigma=1e-3):
    """sin(2 pi f0 (t - t0)) exp(-(t - t0)^2 / (2 sigma^2))."""
    t = np.asarray(t, dtype=np.float64)
    return np.sin(2 * np.pi * f0 * (t - t0)) * np.exp(-((t - t0) ** 2) / (2 * sigma ** 2))


def bandlimited_noise(rng, n_series, n_samples, dt, fmax=4000.0):
    """Unit-RMS white noise low-passed to [0, fmax] by zeroing FFT bins above fmax."""
    x = rng.standard_normal((n_series, n_samples))
    X = np.fft.rfft(x, axis=1)
    f = np.fft.rfftfreq(n_samples, d=dt)
    X[:, f > fmax] = 0
    X[:, 0] = 0
    y = np.fft.irfft(X, n=n_samples, axis=1)
    rms = np.sqrt(np.mean(y * y, axis=1, keepdims=True))
    return y / np.maximum(rms, 1e-30)


def bandlimited_noise_tm(rng, n_samples, n_series, dt, fmax=4000.0):
    """bandlimited_noise's kind of signal, generated time-major ((n_samples, n_series) float32)
    with scipy.fft on every core: for series too large to transpose cheaply (the values differ
    from bandlimited_noise's)."""
    from scipy import fft as sfft
    x = rng.standard_normal((n_samples, n_series), dtype=np.float32)
    X = sfft.rfft(x, axis=0, workers=-1)
    del x
    f = np.fft.rfftfreq(n_samples, d=dt)
    X[f > fmax, :] = 0
    X[0, :] = 0
    y = sfft.irfft(X, n=n_samples, axis=0, workers=-1)
    del X
    rms = np.sqrt((y * y).mean(axis=0, dtype=np.float64))
    y /= np.maximum(rms, 1e-30).astype(np.float32)[None, :]
    return np.ascontiguousarray(y, dtype=np.float32)


def pulsating_sphere(g, dt, radius, r, c):
    """Pressure at distance r of a sphere of radius a whose normal derivative dp/dn = g(t).

    p(r, t) = (a / r) q(t - (r - a)/c) with q' = -(c/a) q - c g (outgoing spherical wave).
    g is sampled at t_i = i dt; q is integrated exactly for piecewise-linear g.
    """
    g = np.asarray(g, dtype=np.float64)
    n = len(g)
    lam = c / radius
    e = math.exp(-lam * dt)
    q = np.zeros(n)
    # exact step for q' = -lam q - c g with g linear on [t_i, t_i+1]
    a1 = (1 - e) / lam
    a2 = (dt - a1) / (lam * dt)
    for i in range(n - 1):
        # integral of e^{-lam (dt - s)} (g_i + (g_{i+1} - g_i) s / dt) ds over one step
        q[i + 1] = e * q[i] - c * (g[i] * a1 + (g[i + 1] - g[i]) * a2)
    tq = np.arange(n) * dt
    delay = (r - radius) / c
    return (radius / r) * np.interp(tq - delay, tq, q, left=0.0, right=0.0)


# ---------------------------------------------------------------------- workloads
@dataclass
class Scene:
    name: str
    res: int
    dl: float
    dt: float
    min_pos: tuple
    c: float
    verts: np.ndarray
    tris: np.ndarray
    steps: int
    listener: tuple
    neumann: np.ndarray = None          # (steps, T) float32, or None if generated lazily
    motion: object = None               # callable step -> verts (float32) for moving scenes
    info: dict = field(default_factory=dict)


def c1_static_sphere(steps: int = 1000, c: float = 343.0, res: int = 32) -> Scene:
    radius = 0.05
    length = 2.0 * (2 * radius)               # gridLengthFactor 2 on the bounding-box length
    h = np.float32(length / res)
    dt = np.float32(float(h) / (math.sqrt(3.0) * c))
    V, F = icosphere(3, radius)
    mn = (-length / 2, -length / 2, -length / 2)
    t = np.arange(steps) * float(dt)
    g = gaussian_sine(t).astype(np.float32)
    neumann = np.repeat(g[:, None], F.shape[0], axis=1)
    lis = tuple(int(0.875 * res) for _ in range(3))
    return Scene("c1_static_sphere", res, float(h), float(dt), mn, c, V, F, steps, lis, neumann,
                 info={"radius": radius, "signal": g})


def c2_moving(steps: int = 4000, c: float = 343.0, res: int = 64, seed: int = SEED) -> Scene:
    rng = np.random.default_rng(seed)
    h = np.float32(0.005)
    length = float(h) * res
    dt = np.float32(float(h) / (math.sqrt(3.0) * c))
    cube_v, cube_f = subdivided_cube(3, side=0.08, center=(-0.07, 0.0, 0.0))
    sph_v0, sph_f = icosphere(4, 0.05)
    sph_center = np.array([0.07, 0.0, 0.0])
    V0 = np.concatenate([cube_v, sph_v0 + sph_center.astype(np.float32)])
    F = np.concatenate([cube_f, sph_f + cube_v.shape[0]])
    n_cube_v = cube_v.shape[0]
    axis = rng.standard_normal(3)
    axis /= np.linalg.norm(axis)
    amp, freq, omega = 0.3 * float(h), 200.0, 2.0

    def verts_at(step: int) -> np.ndarray:
        tt = step * float(dt)
        ang = omega * tt
        K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
        R = np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * (K @ K)
        off = sph_center + np.array([amp * math.sin(2 * math.pi * freq * tt), 0.0, 0.0])
        sv = (sph_v0.astype(np.float64) @ R.T) + off
        out = np.concatenate([cube_v.astype(np.float64), sv]).astype(np.float32)
        return out

    nv = V0.shape[0]
    sel = rng.random(nv) < 0.5
    noise = np.zeros((nv, steps), dtype=np.float64)
    noise[sel] = bandlimited_noise(rng, int(sel.sum()), steps, float(dt))
    tri_series = noise[F].mean(axis=1)          # (T, steps): mean of the 3 vertices
    neumann = np.ascontiguousarray(tri_series.T.astype(np.float32))
    mn = (-length / 2, -length / 2, -length / 2)
    lis = tuple(int(0.875 * res) for _ in range(3))
    return Scene("c2_moving", res, float(h), float(dt), mn, c, verts_at(0), F, steps, lis, neumann,
                 motion=verts_at, info={"n_cube_v": n_cube_v, "axis": axis.tolist()})


def large_static_sphere(steps: int = 500, c: float = 343.0, res: int = 512, seed: int = SEED) -> Scene:
    """One large grid: the C5 scene's object (icosphere level 5, radius 0.05 m, each vertex's
    radius scaled by 1 + U[-0.05, 0.05]) at the C5 cell size h = radius / 32 (128^3 over four
    radii), in a res^3 domain (512^3 = 16 radii wide, 1.6 GB of pressure levels), driven by the
    C1 pulse on every triangle. The FDTD's N^3 stream dominates the step's bytes."""
    rng = np.random.default_rng(seed + 5)
    radius = 0.05
    h = np.float32(4 * radius / 128)
    length = float(h) * res
    dt = np.float32(float(h) / (math.sqrt(3.0) * c))
    V, F = icosphere(5, radius)
    V = (V.astype(np.float64) * (1.0 + rng.uniform(-0.05, 0.05, (V.shape[0], 1)))).astype(np.float32)
    mn = (-length / 2, -length / 2, -length / 2)
    t = np.arange(steps) * float(dt)
    g = gaussian_sine(t).astype(np.float32)
    neumann = np.repeat(g[:, None], F.shape[0], axis=1)
    lis = tuple(int(0.875 * res) for _ in range(3))
    return Scene("large_static_sphere", res, float(h), float(dt), mn, c, V, F, steps, lis, neumann,
                 info={"radius": radius, "signal": g})


def c3_scene(res: int, level: int, seed: int, steps: int = 2000, c: float = 343.0) -> Scene:
    """One scene of the batched configuration: an icosphere of the given level, scaled
    anisotropically by U[0.5, 1.5]^3 (radius 0.05 m before scaling), a res^3 grid over twice its
    bounding box, and a per-scene Gaussian-modulated sine on every triangle (centre frequency
    U[500, 2000] Hz, t0 = 3 sigma with sigma = 1 ms)."""
    rng = np.random.default_rng(SEED + 100 + seed)
    V, F = icosphere(level, 0.05)
    V = (V.astype(np.float64) * rng.uniform(0.5, 1.5, 3)).astype(np.float32)
    lo, hi = V.min(axis=0), V.max(axis=0)
    length = 2.0 * float(np.max(hi - lo))
    h = np.float32(length / res)
    dt = np.float32(float(h) / (math.sqrt(3.0) * c))
    ctr = (lo + hi) / 2.0
    mn = tuple(float(x) for x in (ctr - length / 2))
    t = np.arange(steps) * float(dt)
    g = gaussian_sine(t, f0=float(rng.uniform(500.0, 2000.0)), t0=3e-3, sigma=1e-3).astype(np.float32)
    neumann = np.repeat(g[:, None], F.shape[0], axis=1)
    lis = tuple(int(0.875 * res) for _ in range(3))
    return Scene(f"c3_scene_{res}_L{level}", res, float(h), float(dt), mn, c, V, F, steps, lis, neumann,
                 info={"signal": g})


def c3_batch_sizes(n_scenes: int = 64, seed: int = SEED):
    """(grid edge, icosphere level) of each scene of the batched configuration: edges drawn
    uniformly from 24..96, levels from 1..4 (seeded)."""
    rng = np.random.default_rng(seed + 3)
    res = rng.integers(24, 97, n_scenes)
    lv = rng.integers(1, 5, n_scenes)
    return [(int(r), int(v)) for r, v in zip(res, lv)]


def c3_batch(n_scenes: int = 64, steps: int = 2000, seed: int = SEED):
    """The batched configuration: n_scenes independent scenes of uneven sizes (c3_scene each,
    scene i seeded with i)."""
    return [c3_scene(r, v, seed=i, steps=steps) for i, (r, v) in enumerate(c3_batch_sizes(n_scenes, seed))]


def c4_adjoint(steps: int = 2000, c: float = 343.0, res: int = 48, seed: int = SEED) -> Scene:
    """The adjoint configuration: icosphere L3 (1280 triangles), radius 0.05 m, centred in a res^3
    grid over twice its diameter, listener at relative (0.875, 0.875, 0.875). `neumann` holds the
    inputs the gradient is taken at (seeded band-limited noise, 0-4 kHz, RMS 0.5); info["target"]
    holds the seeded per-triangle accelerations (band-limited noise, unit RMS) whose listener
    signal is the loss's target."""
    rng = np.random.default_rng(seed + 4)
    radius = 0.05
    length = 2.0 * (2 * radius)
    h = np.float32(length / res)
    dt = np.float32(float(h) / (math.sqrt(3.0) * c))
    V, F = icosphere(3, radius)
    mn = (-length / 2, -length / 2, -length / 2)
    target_acc = bandlimited_noise(rng, F.shape[0], steps, float(dt)).T.astype(np.float32)
    g0 = (0.5 * bandlimited_noise(rng, F.shape[0], steps, float(dt))).T.astype(np.float32)
    lis = tuple(int(0.875 * res) for _ in range(3))
    return Scene("c4_adjoint", res, float(h), float(dt), mn, c, V, F, steps, lis, np.ascontiguousarray(g0),
                 info={"radius": radius, "target_acc": np.ascontiguousarray(target_acc)})


def c5_scene(index: int, steps: int = 8192, c: float = 343.0, res: int = 128, seed: int = SEED,
             n_listen: int = 8) -> Scene:
    """One scene of the batched adjoint configuration: icosphere L5 (10242 vertices, 20480
    triangles) of radius 0.05 m with each vertex's radius scaled by 1 + U[-0.05, 0.05] (seeded per
    scene), centred in a res^3 grid over four radii (h = radius / 32 at 128^3), n_listen listeners
    at seeded positions 1.3 to 1.9 radii from the centre. info["target_acc"]: the seeded
    per-triangle accelerations (band-limited noise, 0-4 kHz, unit RMS) whose listener signals
    are the reconstruction target; `neumann` is None (the unknowns start at zero)."""
    rng = np.random.default_rng(seed + 500 + index)
    radius = 0.05
    length = 4 * radius
    h = np.float32(length / res)
    dt = np.float32(float(h) / (math.sqrt(3.0) * c))
    V, F = icosphere(5, radius)
    V = (V.astype(np.float64) * (1.0 + rng.uniform(-0.05, 0.05, (V.shape[0], 1)))).astype(np.float32)
    mn = (-length / 2, -length / 2, -length / 2)
    lis = []
    while len(lis) < n_listen:
        d = rng.standard_normal(3)
        d *= rng.uniform(1.3, 1.9) * radius / np.linalg.norm(d)
        cell = tuple(int(np.floor((x - m) / float(h))) for x, m in zip(d, mn))
        if all(4 <= q <= res - 5 for q in cell) and cell not in lis:
            lis.append(cell)
    target_acc = bandlimited_noise_tm(rng, steps, F.shape[0], float(dt))
    return Scene(f"c5_scene_{index}", res, float(h), float(dt), mn, c, V, F, steps, lis[0], None,
                 info={"listeners": lis, "target_acc": np.ascontiguousarray(target_acc), "radius": radius})


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

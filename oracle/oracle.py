"""ctypes binding of the CPU oracle (liboracle.so) — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module,
as the checker. The product (libpppm_hip.so and its bindings) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF32, FP64 = 0, 1
SINGLE, DOUBLE = 0, 1

LWH_DTYPE = np.dtype([("s", "<u2", (16,)), ("ss", "<f4"), ("d", "<u2", (16,)), ("sd", "<f4")])

_lib = None


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp, i, f, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_double
        sig = {
            "orc_create": ([i, f, f, vp, f, f, i, i], vp),
            "orc_destroy": ([vp], None),
            "orc_set_mesh": ([vp, vp, i, vp, i], i),
            "orc_update_mesh": ([vp, vp], i),
            "orc_step": ([vp, vp], i),
            "orc_fdtd_step": ([vp], i),
            "orc_solve_far": ([vp], i),
            "orc_set_neumann": ([vp, vp], i),
            "orc_update_dirichlet": ([vp], i),
            "orc_solve_near": ([vp], i),
            "orc_solve_far_direct": ([vp], i),
            "orc_solve_near_direct": ([vp], i),
            "orc_time_idx": ([vp], i),
            "orc_set_time_idx": ([vp, i], None),
            "orc_get_grid": ([vp, i, vp], i),
            "orc_set_grid": ([vp, i, vp], i),
            "orc_get_far": ([vp, vp], i),
            "orc_read_cell": ([vp, i, i, i], f),
            "orc_get_history": ([vp, vp, vp], i),
            "orc_set_history": ([vp, vp, vp], i),
            "orc_reset": ([vp], None),
            "orc_stats": ([vp, vp], i),
            "orc_get_triangles": ([vp, vp], i),
            "orc_get_c3": ([vp, vp, vp, vp], i),
            "orc_get_b4": ([vp, vp, vp, vp], i),
            "orc_export_grid_weights": ([vp, vp], i),
            "orc_import_grid_weights": ([vp, vp], i),
            "orc_export_face_weights": ([vp, vp, vp, vp], i),
            "orc_import_face_weights": ([vp, vp], i),
            "orc_get_interp_weights": ([vp, vp], i),
            "orc_f2h": ([f], ctypes.c_ushort),
            "orc_h2f": ([ctypes.c_ushort], f),
            "orc_tdbem_init": ([i, f, f, vp, vp], None),
            "orc_scaled_fft": ([i, vp, d, vp], None),
            "orc_scaled_idft": ([vp, d, vp], None),
            "orc_face2point": ([i, vp, vp, d, d, i, vp], None),
            "orc_face2face": ([i, vp, vp, vp, d, d, i, f, vp], None),
            "orc_laplace_weight_point": ([i, f, f, f, vp, vp, i, vp], None),
            "orc_singular_fp64": ([vp, vp, vp, d, d, i, i, vp], None),
            "orc_classify": ([vp, i], i),
            "orc_get_cells": ([vp, vp, vp, vp, vp, vp, vp], i),
            "orc_set_cells": ([vp, vp, vp, vp, vp], i),
            "orc_linear_forward": ([vp, vp, i, vp, i, vp], i),
            "orc_linear_adjoint": ([vp, vp, i, vp, i, vp], i),
            "orc_linear_fd": ([vp, vp, vp, i, vp, i, i, vp, vp, d, vp], i),
            "orc_fill_fresh": ([vp, vp, vp], i),
            "orc_set_solid_zero": ([vp], i),
            "orc_vertex_normals": ([vp, vp], i),
            "orc_closest_coeff": ([vp, vp, vp], None),
            "orc_det_exp": ([d], d),
            "orc_det_sincos": ([d, vp, vp], None),
            "orc_modal_coeffs": ([vp, i, f, f, f, vp], None),
            "orc_modal_surface_matrix": ([vp, i, i, vp, i, vp, vp], None),
            "orc_modal_iir": ([vp, i, f, vp, vp, i, vp], None),
            "orc_modal_accs": ([vp, i, i, vp, i, vp], None),
            "orc_ffat_collect": ([vp, i, i, vp], None),
            "orc_slice_image": ([vp, i, f, vp, i, i, f, vp, vp, i, vp], None),
            "orc_modal_forces": ([vp, i, vp, vp, vp, vp, i, f, vp], None),
        }
        for name, (a, r) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = a
            fn.restype = r
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """CPU restatement of PPPMSolver (src/algorithm/pppm.h) for parity checks."""

    def __init__(self, res, dl, dt, min_pos=(0, 0, 0), c=343.0, reflect_coeff=0.83, mode=FP64, pml_width=0):
        self.L = lib()
        mn = np.asarray(min_pos, dtype=np.float32)
        self.s = self.L.orc_create(int(res), float(dl), float(dt), _p(mn), float(c), float(reflect_coeff),
                                   int(mode), int(pml_width))
        if not self.s:
            raise ValueError("orc_create failed")
        self.N = int(res)
        self.nt = 0

    def __del__(self):
        try:
            if self.s:
                self.L.orc_destroy(self.s)
        except Exception:
            pass

    def _ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"oracle {what} failed ({rc})")

    def set_mesh(self, verts, tris):
        v = np.ascontiguousarray(verts, dtype=np.float32)
        t = np.ascontiguousarray(tris, dtype=np.int32)
        self._ok(self.L.orc_set_mesh(self.s, _p(v), v.shape[0], _p(t), t.shape[0]), "set_mesh")
        self.nt = t.shape[0]

    def update_mesh(self, verts):
        v = np.ascontiguousarray(verts, dtype=np.float32)
        self._ok(self.L.orc_update_mesh(self.s, _p(v)), "update_mesh")

    def step(self, neumann):
        n = np.ascontiguousarray(neumann, dtype=np.float32)
        self._ok(self.L.orc_step(self.s, _p(n)), "step")

    def fdtd_step(self):
        self._ok(self.L.orc_fdtd_step(self.s), "fdtd_step")

    def solve_far(self):
        self._ok(self.L.orc_solve_far(self.s), "solve_far")

    def set_neumann(self, neumann):
        n = np.ascontiguousarray(neumann, dtype=np.float32)
        self._ok(self.L.orc_set_neumann(self.s, _p(n)), "set_neumann")

    def update_dirichlet(self):
        self._ok(self.L.orc_update_dirichlet(self.s), "update_dirichlet")

    def solve_near(self):
        self._ok(self.L.orc_solve_near(self.s), "solve_near")

    def solve_far_direct(self):
        self._ok(self.L.orc_solve_far_direct(self.s), "solve_far_direct")

    def solve_near_direct(self):
        self._ok(self.L.orc_solve_near_direct(self.s), "solve_near_direct")

    def reset(self):
        self.L.orc_reset(self.s)

    def time_idx(self):
        return self.L.orc_time_idx(self.s)

    def set_time_idx(self, t):
        self.L.orc_set_time_idx(self.s, int(t))

    def grid(self, rel=0):
        out = np.empty(self.N ** 3, dtype=np.float32)
        self.L.orc_get_grid(self.s, int(rel), _p(out))
        return out.reshape(self.N, self.N, self.N)

    def set_grid(self, rel, g):
        g = np.ascontiguousarray(g, dtype=np.float32)
        self.L.orc_set_grid(self.s, int(rel), _p(g))

    def far(self):
        out = np.empty(self.N ** 3, dtype=np.float32)
        self.L.orc_get_far(self.s, _p(out))
        return out.reshape(self.N, self.N, self.N)

    def read(self, x, y, z):
        return self.L.orc_read_cell(self.s, int(x), int(y), int(z))

    def history(self):
        n = np.empty((self.nt, 32), dtype=np.float32)
        d = np.empty((self.nt, 32), dtype=np.float32)
        self.L.orc_get_history(self.s, _p(n), _p(d))
        return n, d

    def set_history(self, n, d):
        n = np.ascontiguousarray(n, dtype=np.float32)
        d = np.ascontiguousarray(d, dtype=np.float32)
        self.L.orc_set_history(self.s, _p(n), _p(d))

    def stats(self):
        out = np.zeros(8, dtype=np.int64)
        self.L.orc_stats(self.s, _p(out))
        keys = ["T", "C3", "P3", "B4", "P4", "grid_recomputes", "face_recomputes"]
        return dict(zip(keys, (int(x) for x in out[:7])))

    def triangles(self):
        out = np.zeros((self.nt, 13), dtype=np.float32)
        self.L.orc_get_triangles(self.s, _p(out))
        return out

    def c3(self):
        st = self.stats()
        cells = np.zeros(max(st["C3"], 1), np.int32)
        start = np.zeros(st["C3"] + 1, np.int32)
        pairs = np.zeros(max(st["P3"], 1), np.int32)
        self.L.orc_get_c3(self.s, _p(cells), _p(start), _p(pairs))
        return cells[:st["C3"]], start, pairs[:st["P3"]]

    def b4(self):
        st = self.stats()
        cells = np.zeros(max(st["B4"], 1), np.int32)
        start = np.zeros(st["B4"] + 1, np.int32)
        self.L.orc_get_b4(self.s, _p(cells), _p(start), None)
        tris = np.zeros(max(int(start[-1]), 1), np.int32)
        self.L.orc_get_b4(self.s, None, None, _p(tris))
        return cells[:st["B4"]], start, tris[:int(start[-1])]

    def grid_weights(self):
        rec = np.zeros(self.nt * 27 * 2, dtype=LWH_DTYPE)
        self.L.orc_export_grid_weights(self.s, _p(rec))
        return rec

    def set_grid_weights(self, rec):
        rec = np.ascontiguousarray(rec, dtype=LWH_DTYPE)
        self.L.orc_import_grid_weights(self.s, _p(rec))

    def face_weights(self):
        st = self.stats()
        start = np.zeros(self.nt + 1, np.int32)
        nbr = np.zeros(max(st["P4"], 1), np.int32)
        rec = np.zeros(max(st["P4"], 1), dtype=LWH_DTYPE)
        self.L.orc_export_face_weights(self.s, _p(start), _p(nbr), _p(rec))
        return start, nbr[:st["P4"]], rec[:st["P4"]]

    def set_face_weights(self, rec):
        rec = np.ascontiguousarray(rec, dtype=LWH_DTYPE)
        self.L.orc_import_face_weights(self.s, _p(rec))

    def interp_weights(self):
        out = np.zeros((self.nt, 8), np.float32)
        self.L.orc_get_interp_weights(self.s, _p(out))
        return out

    # cell classification / fresh cells (ghost-cell solver)
    def classify(self, thin_shell=False):
        r = self.L.orc_classify(self.s, int(bool(thin_shell)))
        if r < 0:
            raise RuntimeError("oracle classify failed")
        return r

    def cells(self):
        NN = self.N ** 3
        out = {"type": np.zeros(NN, np.int32), "ghost_idx": np.zeros(NN, np.int32),
               "nearest": np.zeros(NN, np.int32), "point": np.zeros((NN, 3), np.float32),
               "reflect": np.zeros((NN, 3), np.float32), "dist": np.zeros(NN, np.float32)}
        self._ok(self.L.orc_get_cells(self.s, *(_p(out[k]) for k in
                                                 ("type", "ghost_idx", "nearest", "point", "reflect", "dist"))),
                 "get_cells")
        return out

    # the step as a linear map in FP64 (adjoint_oracle.cpp)
    def _cells(self, listeners):
        lx = np.asarray(listeners, dtype=np.int64).reshape(-1, 3)
        return np.ascontiguousarray((lx[:, 0] * self.N + lx[:, 1]) * self.N + lx[:, 2], dtype=np.int32)

    def linear_forward(self, g, listeners):
        """FP64 forward from the zero state: g (T, nt) -> listener samples (T, n_listen)."""
        g = np.ascontiguousarray(g, dtype=np.float64)
        lc = self._cells(listeners)
        out = np.zeros((g.shape[0], len(lc)), np.float64)
        self._ok(self.L.orc_linear_forward(self.s, _p(g), g.shape[0], _p(lc), len(lc), _p(out)), "linear_forward")
        return out

    def linear_adjoint(self, gout, listeners):
        """FP64 adjoint: listener cotangents (T, n_listen) -> gradient of the inputs (T, nt)."""
        gout = np.ascontiguousarray(gout, dtype=np.float64)
        lc = self._cells(listeners)
        grad = np.zeros((gout.shape[0], self.nt), np.float64)
        self._ok(self.L.orc_linear_adjoint(self.s, _p(gout), gout.shape[0], _p(lc), len(lc), _p(grad)),
                 "linear_adjoint")
        return grad

    def linear_fd(self, g, target, listeners, steps, tris, eps=1e-3):
        """Central differences of sum (out - target)^2 at the entries (steps[e], tris[e])."""
        g = np.ascontiguousarray(g, dtype=np.float64)
        target = np.ascontiguousarray(target, dtype=np.float64)
        lc = self._cells(listeners)
        st = np.ascontiguousarray(steps, dtype=np.int32)
        tr = np.ascontiguousarray(tris, dtype=np.int32)
        fd = np.zeros(len(st), np.float64)
        self._ok(self.L.orc_linear_fd(self.s, _p(g), _p(target), g.shape[0], _p(lc), len(lc), len(st), _p(st),
                                      _p(tr), float(eps), _p(fd)), "linear_fd")
        return fd

    def set_cells(self, ctype, ctype_old, nearest, point3):
        """Test hook: cell types now / before, nearest triangles and points (N^3 entries)."""
        a = np.ascontiguousarray(ctype, dtype=np.int32)
        b = np.ascontiguousarray(ctype_old, dtype=np.int32)
        n = np.ascontiguousarray(nearest, dtype=np.int32)
        p = np.ascontiguousarray(point3, dtype=np.float32)
        self._ok(self.L.orc_set_cells(self.s, _p(a), _p(b), _p(n), _p(p)), "set_cells")

    def fill_fresh(self, acc_now, acc_old):
        a = np.ascontiguousarray(acc_now, np.float32)
        b = np.ascontiguousarray(acc_old, np.float32)
        return int(self.L.orc_fill_fresh(self.s, _p(a), _p(b)))

    def set_solid_zero(self):
        self._ok(self.L.orc_set_solid_zero(self.s), "set_solid_zero")

    def vertex_normals(self, nv):
        out = np.zeros((nv, 3), np.float32)
        self._ok(self.L.orc_vertex_normals(self.s, _p(out)), "vertex_normals")
        return out


# ---------------------------------------------------------------------- primitives
def f2h(x: float) -> int:
    return int(lib().orc_f2h(float(x)))


def h2f(h: int) -> float:
    return float(lib().orc_h2f(int(h)))


def tdbem_init(mode, dt, c):
    lam = ctypes.c_double()
    wave = np.zeros(32, np.float64)
    lib().orc_tdbem_init(int(mode), float(dt), float(c), ctypes.byref(lam), _p(wave))
    return lam.value, wave[0::2] + 1j * wave[1::2]


def scaled_fft(mode, v16, lam):
    v = np.zeros(32, np.float64)
    v[0::2] = np.real(v16)
    v[1::2] = np.imag(v16)
    out = np.zeros(16, np.float64)
    lib().orc_scaled_fft(int(mode), _p(v), float(lam), _p(out))
    return out


def scaled_idft(v16, lam):
    v = np.zeros(32, np.float64)
    v[0::2] = np.real(v16)
    v[1::2] = np.imag(v16)
    out = np.zeros(16, np.float64)
    lib().orc_scaled_idft(_p(v), float(lam), _p(out))
    return out


def face2point(mode, tri9, pt3, k, layer):
    tri = np.ascontiguousarray(tri9, np.float32).reshape(9)
    pt = np.ascontiguousarray(pt3, np.float32).reshape(3)
    out = np.zeros(2, np.float64)
    lib().orc_face2point(int(mode), _p(tri), _p(pt), float(np.real(k)), float(np.imag(k)), int(layer), _p(out))
    return out[0] + 1j * out[1]


def face2face(mode, verts, src3, trg3, k, layer, threshold):
    v = np.ascontiguousarray(verts, np.float32).reshape(-1)
    s = np.ascontiguousarray(src3, np.int32)
    t = np.ascontiguousarray(trg3, np.int32)
    out = np.zeros(2, np.float64)
    lib().orc_face2face(int(mode), _p(v), _p(s), _p(t), float(np.real(k)), float(np.imag(k)), int(layer),
                        float(threshold), _p(out))
    return out[0] + 1j * out[1]


def singular_fp64(verts, src3, trg3, k, layer, gauss_n):
    v = np.ascontiguousarray(verts, np.float32).reshape(-1)
    s = np.ascontiguousarray(src3, np.int32)
    t = np.ascontiguousarray(trg3, np.int32)
    out = np.zeros(2, np.float64)
    lib().orc_singular_fp64(_p(v), _p(s), _p(t), float(np.real(k)), float(np.imag(k)), int(layer), int(gauss_n),
                            _p(out))
    return out[0] + 1j * out[1]


def laplace_weight_point(mode, dt, c, h, tri9, pt3, layer):
    tri = np.ascontiguousarray(tri9, np.float32).reshape(9)
    pt = np.ascontiguousarray(pt3, np.float32).reshape(3)
    out = np.zeros(16, np.float64)
    lib().orc_laplace_weight_point(int(mode), float(dt), float(c), float(h), _p(tri), _p(pt), int(layer), _p(out))
    return out


def closest_coeff(p3, tri9):
    p = np.ascontiguousarray(p3, np.float32).reshape(3)
    t = np.ascontiguousarray(tri9, np.float32).reshape(9)
    out = np.zeros(3, np.float32)
    lib().orc_closest_coeff(_p(p), _p(t), _p(out))
    return out


def modal_coeffs(eigvals, timestep, alpha, beta):
    """ModalInfo::SetCoeffs for every mode -> (M, 3) float32 (coeff1, coeff2, coeff3)."""
    ev = np.ascontiguousarray(eigvals, np.float32)
    out = np.zeros((ev.size, 3), np.float32)
    lib().orc_modal_coeffs(_p(ev), ev.size, ctypes.c_float(timestep), ctypes.c_float(alpha), ctypes.c_float(beta),
                           _p(out))
    return out


def modal_surface_matrix(eig, tris, normals):
    """MatrixVertToTri: eig (3 nv, M) float32, tris (T, 3), normals (T, 3) -> (T, M) float32."""
    e = np.ascontiguousarray(eig, np.float32)
    t = np.ascontiguousarray(tris, np.int32)
    n = np.ascontiguousarray(normals, np.float32)
    out = np.zeros((t.shape[0], e.shape[1]), np.float32)
    lib().orc_modal_surface_matrix(_p(e), e.shape[0] // 3, e.shape[1], _p(t), t.shape[0], _p(n), _p(out))
    return out


def modal_iir(c123, timestep, state, forces):
    """IIR recursion: state (M, 3) updated in place, forces (S, M) -> Q (S, M) float32."""
    c = np.ascontiguousarray(c123, np.float32)
    f = np.ascontiguousarray(forces, np.float32)
    assert state.dtype == np.float32 and state.flags.c_contiguous
    q = np.zeros_like(f)
    lib().orc_modal_iir(_p(c), c.shape[0], ctypes.c_float(timestep), _p(state), _p(f), f.shape[0], _p(q))
    return q


def modal_accs(msurf, q):
    """acc[s][t] = sum_i M[t][i] Q[s][i] in mode order -> (S, T) float32."""
    m = np.ascontiguousarray(msurf, np.float32)
    qq = np.ascontiguousarray(q, np.float32)
    out = np.zeros((qq.shape[0], m.shape[0]), np.float32)
    lib().orc_modal_accs(_p(m), m.shape[0], m.shape[1], _p(qq), qq.shape[0], _p(out))
    return out


def ffat_collect(grid, w, fmap):
    """collect_ffat_map: fmap (6, 2w, 2w) float32 += |p| on the six faces of the cube of half
    width w around the grid centre (in place)."""
    g = np.ascontiguousarray(grid, np.float32)
    assert fmap.dtype == np.float32 and fmap.flags.c_contiguous and fmap.shape == (6, 2 * w, 2 * w)
    lib().orc_ffat_collect(_p(g), g.shape[0], int(w), _p(fmap))
    return fmap


def slice_image(grid, h, min_pos, axis, index, vmax, verts=None, tris=None):
    """save_grid's picture of plane `index` normal to `axis`: (W, W, 4) uint8 in PNG row order."""
    g = np.ascontiguousarray(grid, np.float32)
    N = g.shape[0]
    W = N * (1080 // N)
    out = np.zeros((W, W, 4), np.uint8)
    mn = np.ascontiguousarray(min_pos, np.float32)
    if verts is None:
        v = np.zeros((1, 3), np.float32)
        t = np.zeros((1, 3), np.int32)
        nt = 0
    else:
        v = np.ascontiguousarray(verts, np.float32)
        t = np.ascontiguousarray(tris, np.int32)
        nt = t.shape[0]
    lib().orc_slice_image(_p(g), N, float(h), _p(mn), int(axis), int(index), float(vmax), _p(v), _p(t), nt, _p(out))
    return out


def modal_forces(eig, imp_time, imp_vert, imp_vec, active, time):
    """Modal forces of the active impulses at `time` -> (M,) float32."""
    e = np.ascontiguousarray(eig, np.float32)
    it = np.ascontiguousarray(imp_time, np.float32)
    iv = np.ascontiguousarray(imp_vert, np.int32)
    ivec = np.ascontiguousarray(imp_vec, np.float32)
    act = np.ascontiguousarray(active, np.int32)
    out = np.zeros(e.shape[1], np.float32)
    lib().orc_modal_forces(_p(e), e.shape[1], _p(it), _p(iv), _p(ivec), _p(act), it.size, ctypes.c_float(time),
                           _p(out))
    return out


def det_exp(x):
    return float(lib().orc_det_exp(float(x)))


def det_sincos(x):
    s, c = ctypes.c_double(), ctypes.c_double()
    lib().orc_det_sincos(float(x), ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value

"""Python binding of libpppm_hip.so (ctypes over the C ABI in include/pppm_hip.h).

`PPPMSolver` mirrors the reference's solver interface (src/algorithm/pppm.h:11-164):
set_mesh / update_mesh / update_grid_and_face (alias update_step) / solve_fdtd_far /
set_neumann_condition / update_dirichlet / solve_fdtd_near / dt / dl / res / get_grid /
get_triangles / clear, plus `run` (many steps with on-device listener recording).

Arrays may be numpy arrays (host) or torch tensors on the GPU (device pointers are
passed through). The library must be built (``make -C pppm-tdbem_amd``); there is no
fallback: a missing library raises at import of the solver.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpppm_hip.so")

PPPM_FLAG_HONOR_PML = 1

EXPORTS = [
    "pppm_create", "pppm_destroy", "pppm_set_stream", "pppm_last_error", "pppm_set_mesh",
    "pppm_update_mesh", "pppm_step", "pppm_update_mesh_and_step", "pppm_run", "pppm_run_noalloc", "pppm_fdtd_step",
    "pppm_solve_far",
    "pppm_set_neumann", "pppm_update_dirichlet", "pppm_solve_near", "pppm_reset",
    "pppm_time_idx", "pppm_grid_ptr", "pppm_read_cells", "pppm_history_ptrs", "pppm_triangles",
    "pppm_stats", "pppm_get_grid_weights", "pppm_set_grid_weights", "pppm_get_face_weights",
    "pppm_set_face_weights", "pppm_get_interp_weights", "pppm_get_c3", "pppm_get_b4",
    "pppm_set_timing", "pppm_get_timing", "pppm_get_update_work", "pppm_run_moving", "pppm_classify", "pppm_fill_fresh",
    "pppm_set_solid_zero", "pppm_get_cells", "pppm_set_time_idx", "pppm_ffat_collect",
    "pppm_ghost_setup", "pppm_ghost_set_neumann", "pppm_ghost_solve", "pppm_ghost_step", "pppm_ghost_update_mesh",
    "pppm_ghost_get", "pppm_slice_image",
    "pppm_modal_create", "pppm_modal_destroy", "pppm_modal_run", "pppm_modal_state", "pppm_modal_set_state",
    "pppm_modal_surface_matrix", "pppm_modal_coeffs",
    "pppm_batch_create", "pppm_batch_destroy", "pppm_batch_last_error", "pppm_batch_set_stream", "pppm_batch_run",
    "pppm_batch_autotune", "pppm_batch_variants", "pppm_batch_set_variants", "pppm_batch_set_timing",
    "pppm_batch_get_timing",
    "pppm_adjoint_reset", "pppm_adjoint_run", "pppm_state_bytes", "pppm_save_state", "pppm_load_state",
    "pppm_batch_adjoint_reset", "pppm_batch_adjoint_run",
]

# LayerWeightHalf record (src/algorithm/bem.h:163-169): 16 x fp16, f32, 16 x fp16, f32
LWH_DTYPE = np.dtype([("s", "<u2", (16,)), ("ss", "<f4"), ("d", "<u2", (16,)), ("sd", "<f4")])
assert LWH_DTYPE.itemsize == 72

_lib = None


class PPPMError(RuntimeError):
    pass


def load_library(path: str = LIB_PATH):
    """Loads libpppm_hip.so and declares the C signatures. Raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise PPPMError(f"{path} is not built (run `make -C {_HERE}`)")
    lib = ctypes.CDLL(path)
    vp, i, f, ll = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_longlong
    P = ctypes.POINTER
    sig = {
        "pppm_create": ([P(vp), i, f, f, vp, i, f, f, ctypes.c_uint], i),
        "pppm_destroy": ([vp], i),
        "pppm_set_stream": ([vp, vp], i),
        "pppm_last_error": ([vp], ctypes.c_char_p),
        "pppm_set_mesh": ([vp, vp, i, vp, i, i], i),
        "pppm_update_mesh": ([vp, vp, i], i),
        "pppm_step": ([vp, vp, i], i),
        "pppm_update_mesh_and_step": ([vp, vp, vp, i], i),
        "pppm_run": ([vp, vp, ll, i, vp, i, vp, i], i),
        "pppm_run_noalloc": ([vp, vp, ll, i, vp, i, vp], i),
        "pppm_fdtd_step": ([vp], i),
        "pppm_solve_far": ([vp], i),
        "pppm_set_neumann": ([vp, vp, i], i),
        "pppm_update_dirichlet": ([vp], i),
        "pppm_solve_near": ([vp], i),
        "pppm_reset": ([vp], i),
        "pppm_time_idx": ([vp, P(i)], i),
        "pppm_grid_ptr": ([vp, i, P(vp)], i),
        "pppm_read_cells": ([vp, vp, i, vp], i),
        "pppm_ffat_collect": ([vp, i, vp], i),
        "pppm_history_ptrs": ([vp, P(vp), P(vp)], i),
        "pppm_triangles": ([vp, vp], i),
        "pppm_stats": ([vp, vp], i),
        "pppm_get_grid_weights": ([vp, vp], i),
        "pppm_set_grid_weights": ([vp, vp], i),
        "pppm_get_face_weights": ([vp, vp, vp, vp], i),
        "pppm_set_face_weights": ([vp, vp], i),
        "pppm_get_interp_weights": ([vp, vp], i),
        "pppm_get_c3": ([vp, vp, vp, vp], i),
        "pppm_get_b4": ([vp, vp, vp, vp], i),
        "pppm_set_timing": ([vp, i], i),
        "pppm_get_timing": ([vp, vp, vp], i),
        "pppm_get_update_work": ([vp, vp], i),
        "pppm_run_moving": ([vp, vp, ll, vp, ll, i, vp, i, vp, i], i),
        "pppm_classify": ([vp, i, P(i)], i),
        "pppm_fill_fresh": ([vp, vp, vp, i, P(i)], i),
        "pppm_set_solid_zero": ([vp], i),
        "pppm_get_cells": ([vp, vp, vp, vp, vp, vp, vp], i),
        "pppm_set_time_idx": ([vp, i], i),
        "pppm_ghost_setup": ([vp, f, P(i)], i),
        "pppm_ghost_set_neumann": ([vp, vp, i], i),
        "pppm_ghost_solve": ([vp, P(i)], i),
        "pppm_ghost_step": ([vp, vp, i], i),
        "pppm_ghost_update_mesh": ([vp, vp, i, P(i)], i),
        "pppm_ghost_get": ([vp, vp, vp, vp, vp, vp, vp, vp, vp], i),
        "pppm_slice_image": ([vp, i, i, f, vp, P(i)], i),
        "pppm_modal_create": ([P(vp), vp, i, i, vp, vp, i, vp, f, f, f, vp], i),
        "pppm_modal_destroy": ([vp], i),
        "pppm_modal_run": ([vp, vp, i, vp, i], i),
        "pppm_modal_state": ([vp, vp], i),
        "pppm_modal_set_state": ([vp, vp], i),
        "pppm_modal_surface_matrix": ([vp, vp], i),
        "pppm_modal_coeffs": ([vp, vp], i),
        "pppm_batch_create": ([P(vp), vp, i], i),
        "pppm_batch_destroy": ([vp], i),
        "pppm_batch_last_error": ([vp], ctypes.c_char_p),
        "pppm_batch_set_stream": ([vp, vp], i),
        "pppm_batch_run": ([vp, vp, vp, i, i, vp, vp, i], i),
        "pppm_batch_autotune": ([vp, i, vp], i),
        "pppm_batch_variants": ([vp, vp], i),
        "pppm_batch_set_variants": ([vp, vp], i),
        "pppm_batch_set_timing": ([vp, i], i),
        "pppm_batch_get_timing": ([vp, vp, vp], i),
        "pppm_adjoint_reset": ([vp, i], i),
        "pppm_batch_adjoint_reset": ([vp, i], i),
        "pppm_batch_adjoint_run": ([vp, vp, i, i, vp, vp, vp, i], i),
        "pppm_adjoint_run": ([vp, vp, i, vp, i, vp, ll, i], i),
        "pppm_state_bytes": ([vp, P(ll)], i),
        "pppm_save_state": ([vp, vp], i),
        "pppm_load_state": ([vp, vp], i),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


_NP = {"f32": np.float32, "i32": np.int32, "u8": np.uint8}


class _Arg:
    """A validated array argument: `ptr` for the C call, `dev` (1: device memory), and `keep`,
    the host array the pointer refers to (held for the duration of the call)."""

    __slots__ = ("ptr", "dev", "keep")

    def __init__(self, ptr, dev, keep):
        self.ptr, self.dev, self.keep = ptr, dev, keep


def _arg(a, name, count, kind="f32", where="any"):
    """Checks an array argument before it reaches the C ABI: element type (`kind`: f32 / i32),
    at least `count` elements, and its memory (`where`: "device", "host" or "any"). Host data
    may be any array-like (converted to a contiguous numpy array of the right type); a torch
    tensor must already have the right dtype and be contiguous. Raises PPPMError, so a short or
    host buffer never reaches a kernel as a device pointer."""
    if a is None:
        raise PPPMError(f"{name} is required")
    if hasattr(a, "data_ptr"):  # torch tensor
        import torch
        want = {"f32": torch.float32, "i32": torch.int32, "u8": torch.uint8}[kind]
        if a.dtype != want:
            raise PPPMError(f"{name} must be {want}, got {a.dtype}")
        if not a.is_contiguous():
            raise PPPMError(f"{name} must be contiguous")
        dev = int(a.is_cuda)
        if a.numel() < count:
            raise PPPMError(f"{name} holds {a.numel()} elements, needs {count}")
        keep = a
        ptr = ctypes.c_void_p(a.data_ptr())
    else:
        arr = np.asarray(a)
        if arr.dtype != _NP[kind]:
            if kind == "i32" and not np.issubdtype(arr.dtype, np.integer):
                raise PPPMError(f"{name} must be an integer array")
            arr = arr.astype(_NP[kind])
        arr = np.ascontiguousarray(arr)
        if arr.size < count:
            raise PPPMError(f"{name} holds {arr.size} elements, needs {count}")
        dev, keep, ptr = 0, arr, arr.ctypes.data_as(ctypes.c_void_p)
    if where == "device" and not dev:
        raise PPPMError(f"{name} must be device memory (a CUDA tensor)")
    if where == "host" and dev:
        raise PPPMError(f"{name} must be host memory")
    return _Arg(ptr, dev, keep)


class PPPMSolver:
    """MI355X PPPM solver; the constructor mirrors PPPMSolver(res, dl, dt, min_pos, pml_width,
    reflect_coeff) (src/algorithm/pppm.h:39-55) with the speed of sound as a parameter."""

    def __init__(self, res, dl, dt, min_pos=(0.0, 0.0, 0.0), pml_width=0, reflect_coeff=0.83,
                 c=343.0, honor_pml=False):
        self.lib = load_library()
        h = ctypes.c_void_p()
        mn = np.asarray(min_pos, dtype=np.float32)
        self._mn = mn
        flags = PPPM_FLAG_HONOR_PML if honor_pml else 0
        rc = self.lib.pppm_create(ctypes.byref(h), int(res), float(dl), float(dt),
                                  mn.ctypes.data_as(ctypes.c_void_p), int(pml_width),
                                  float(reflect_coeff), float(c), flags)
        if rc != 0:
            raise PPPMError(f"pppm_create failed ({rc})")
        self.h = h
        self._res, self._dl, self._dt, self._c = int(res), float(dl), float(dt), float(c)
        self.min_pos = tuple(float(x) for x in mn)
        self.nt = 0
        self.nv = 0

    # ------------------------------------------------------------------ plumbing
    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.pppm_last_error(self.h)
            raise PPPMError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def clear(self):
        if getattr(self, "h", None):
            self.lib.pppm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.clear()
        except Exception:
            pass

    def set_stream(self, stream_handle: int):
        self._check(self.lib.pppm_set_stream(self.h, ctypes.c_void_p(stream_handle)), "set_stream")

    # ------------------------------------------------------------------ reference-shaped API
    def dt(self):
        return self._dt

    def dl(self):
        return self._dl

    def res(self):
        return self._res

    def min_coord(self):
        return self.min_pos

    def time_idx(self):
        t = ctypes.c_int()
        self._check(self.lib.pppm_time_idx(self.h, ctypes.byref(t)), "time_idx")
        return t.value

    def set_time_idx(self, t):
        self._check(self.lib.pppm_set_time_idx(self.h, int(t)), "set_time_idx")

    def set_mesh(self, verts, tris):
        nv = int(np.shape(verts)[0]) if np.ndim(verts) == 2 else int(np.size(verts)) // 3
        nt = int(np.shape(tris)[0]) if np.ndim(tris) == 2 else int(np.size(tris)) // 3
        v = _arg(verts, "verts", 3 * nv, "f32")
        t = _arg(tris, "tris", 3 * nt, "i32")
        if v.dev != t.dev:
            raise PPPMError("verts and tris must both be host or both device")
        self._check(self.lib.pppm_set_mesh(self.h, v.ptr, nv, t.ptr, nt, v.dev), "set_mesh")
        self.nv, self.nt = nv, nt

    def update_mesh(self, verts):
        v = _arg(verts, "verts", 3 * self.nv, "f32")
        self._check(self.lib.pppm_update_mesh(self.h, v.ptr, v.dev), "update_mesh")

    def update_mesh_and_step(self, verts, neumann):
        """update_mesh(verts) then update_grid_and_face(neumann) in one call; both numpy arrays
        or both device tensors."""
        v = _arg(verts, "verts", 3 * self.nv, "f32")
        n = _arg(neumann, "neumann", self.nt, "f32")
        if v.dev != n.dev:
            raise PPPMError("verts and neumann must both be host or both device")
        self._check(self.lib.pppm_update_mesh_and_step(self.h, v.ptr, n.ptr, v.dev), "update_mesh_and_step")

    def update_grid_and_face(self, neumann):
        n = _arg(neumann, "neumann", self.nt, "f32")
        self._check(self.lib.pppm_step(self.h, n.ptr, n.dev), "step")

    update_step = update_grid_and_face

    def _listeners(self, listeners, out_dev, steps):
        lx = np.ascontiguousarray(np.asarray(listeners, dtype=np.int32).reshape(-1, 3))
        if lx.shape[0] and (lx.min() < 0 or lx.max() >= self._res):
            raise PPPMError("listener outside the grid")
        out = _arg(out_dev, "out", int(steps) * lx.shape[0], "f32", "device") if lx.shape[0] else None
        return lx, out

    def run(self, neumann_series_dev, steps, listeners, out_dev, use_graph=True, stride=None):
        """steps x stride device Neumann series -> out_dev[steps, n_listen] (device)."""
        stride = self.nt if stride is None else int(stride)
        if stride < self.nt or steps < 0:
            raise PPPMError("run: stride must be >= the triangle count and steps >= 0")
        lx, out = self._listeners(listeners, out_dev, steps)
        need = (int(steps) - 1) * stride + self.nt if steps > 0 else 0
        n = _arg(neumann_series_dev, "neumann_series", need, "f32", "device")
        self._check(self.lib.pppm_run(self.h, n.ptr, stride, int(steps), lx.ctypes.data_as(ctypes.c_void_p),
                                      lx.shape[0], out.ptr if out else None, int(use_graph)), "run")

    def run_noalloc(self, neumann_series_dev, steps, listeners, out_dev, stride=None):
        """run() without a graph and without device allocation (pppm_run_noalloc): the gathers'
        per-step buffers come from the library's per-device scratch array. Asynchronous on the
        solver's stream."""
        stride = self.nt if stride is None else int(stride)
        if stride < self.nt or steps < 0:
            raise PPPMError("run_noalloc: stride must be >= the triangle count and steps >= 0")
        lx, out = self._listeners(listeners, out_dev, steps)
        need = (int(steps) - 1) * stride + self.nt if steps > 0 else 0
        n = _arg(neumann_series_dev, "neumann_series", need, "f32", "device")
        self._check(self.lib.pppm_run_noalloc(self.h, n.ptr, stride, int(steps), lx.ctypes.data_as(ctypes.c_void_p),
                                              lx.shape[0], out.ptr if out else None), "run_noalloc")

    def run_moving(self, verts_series_dev, neumann_series_dev, steps, listeners, out_dev, use_graph=True):
        """Moving mesh: before step i the vertices become verts_series_dev[i] (device, steps x nv x 3)."""
        if steps < 0:
            raise PPPMError("run_moving: steps >= 0")
        lx, out = self._listeners(listeners, out_dev, steps)
        v = _arg(verts_series_dev, "verts_series", int(steps) * 3 * self.nv, "f32", "device")
        n = _arg(neumann_series_dev, "neumann_series", int(steps) * self.nt, "f32", "device")
        self._check(self.lib.pppm_run_moving(self.h, v.ptr, int(self.nv * 3), n.ptr, int(self.nt), int(steps),
                                             lx.ctypes.data_as(ctypes.c_void_p), lx.shape[0],
                                             out.ptr if out else None, int(use_graph)), "run_moving")

    # ------------------------------------------------------------------ adjoint and checkpoints
    def adjoint_reset(self, t_last):
        """Zero adjoint state; the next reverse step is forward step t_last."""
        self._check(self.lib.pppm_adjoint_reset(self.h, int(t_last)), "adjoint_reset")

    def adjoint_run(self, gout_dev, steps, listeners, grad_dev, use_graph=True, grad_stride=None):
        """Reverses `steps` steps: gout_dev (device, steps x n_listen, forward time order) ->
        grad_dev (device, steps x grad_stride: the gradient of each step's Neumann input)."""
        lx = np.ascontiguousarray(np.asarray(listeners, dtype=np.int32).reshape(-1, 3))
        if lx.shape[0] and (lx.min() < 0 or lx.max() >= self._res):
            raise PPPMError("listener outside the grid")
        gs = self.nt if grad_stride is None else int(grad_stride)
        if gs < self.nt or steps < 0:
            raise PPPMError("adjoint_run: grad_stride >= triangle count, steps >= 0")
        go = _arg(gout_dev, "gout", int(steps) * lx.shape[0], "f32", "device") if lx.shape[0] else None
        gr = _arg(grad_dev, "grad", (int(steps) - 1) * gs + self.nt if steps > 0 else 0, "f32", "device")
        self._check(self.lib.pppm_adjoint_run(self.h, go.ptr if go else None, int(steps),
                                              lx.ctypes.data_as(ctypes.c_void_p), lx.shape[0], gr.ptr, gs,
                                              int(use_graph)), "adjoint_run")

    def state_bytes(self):
        n = ctypes.c_longlong()
        self._check(self.lib.pppm_state_bytes(self.h, ctypes.byref(n)), "state_bytes")
        return n.value

    def save_state(self, dst_dev):
        """Checkpoint of the forward state into a device tensor of state_bytes() bytes."""
        d = _arg(dst_dev, "state", 0, "u8", "device")
        if dst_dev.numel() * dst_dev.element_size() < self.state_bytes():
            raise PPPMError("state buffer too small")
        self._check(self.lib.pppm_save_state(self.h, d.ptr), "save_state")

    def load_state(self, src_dev):
        d = _arg(src_dev, "state", 0, "u8", "device")
        if src_dev.numel() * src_dev.element_size() < self.state_bytes():
            raise PPPMError("state buffer too small")
        self._check(self.lib.pppm_load_state(self.h, d.ptr), "load_state")

    def fdtd_step(self):
        self._check(self.lib.pppm_fdtd_step(self.h), "fdtd_step")

    def solve_fdtd_far(self):
        self._check(self.lib.pppm_solve_far(self.h), "solve_far")

    def set_neumann_condition(self, neumann):
        n = _arg(neumann, "neumann", self.nt, "f32")
        self._check(self.lib.pppm_set_neumann(self.h, n.ptr, n.dev), "set_neumann")

    def update_dirichlet(self):
        self._check(self.lib.pppm_update_dirichlet(self.h), "update_dirichlet")

    def solve_fdtd_near(self):
        self._check(self.lib.pppm_solve_near(self.h), "solve_near")

    PHASES = ("fdtd", "far", "dirichlet", "near", "update_mesh", "grid_weights", "face_cache", "lists")

    def set_timing(self, enable=True):
        self._check(self.lib.pppm_set_timing(self.h, int(enable)), "set_timing")

    def get_timing(self):
        """{phase: (total_ms, launches)} accumulated since the last call (HIP events)."""
        ms = np.zeros(8, np.float64)
        n = np.zeros(8, np.int64)
        self._check(self.lib.pppm_get_timing(self.h, ms.ctypes.data_as(ctypes.c_void_p),
                                             n.ctypes.data_as(ctypes.c_void_p)), "get_timing")
        return {name: (float(ms[i]), int(n[i])) for i, name in enumerate(self.PHASES)}

    def get_update_work(self):
        """{updates, grid_triangles, face_pairs}: mesh updates completed while timing was on and
        the cache entries they recomputed, since the last call."""
        out = np.zeros(3, np.int64)
        self._check(self.lib.pppm_get_update_work(self.h, out.ctypes.data_as(ctypes.c_void_p)), "get_update_work")
        return {"updates": int(out[0]), "grid_triangles": int(out[1]), "face_pairs": int(out[2])}

    def reset(self):
        self._check(self.lib.pppm_reset(self.h), "reset")

    def grid_ptr(self, rel=0):
        p = ctypes.c_void_p()
        self._check(self.lib.pppm_grid_ptr(self.h, int(rel), ctypes.byref(p)), "grid_ptr")
        return p.value

    def get_grid(self, rel=0):
        """Host copy of grids[t + rel] as an (N, N, N) float32 array (x, y, z; z contiguous)."""
        n = self._res ** 3
        out = np.empty(n, dtype=np.float32)
        rc = _hip_runtime().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.grid_ptr(rel)),
                                      ctypes.c_size_t(4 * n), 2)  # hipMemcpyDeviceToHost (syncs the handle's stream)
        if rc != 0:
            raise PPPMError(f"hipMemcpy failed ({rc})")
        return out.reshape(self._res, self._res, self._res)

    def set_grid(self, rel, values):
        """Overwrites grids[t + rel] from a host array of N^3 floats."""
        a = np.ascontiguousarray(values, dtype=np.float32).reshape(-1)
        if a.size != self._res ** 3:
            raise PPPMError("set_grid: wrong size")
        rc = _hip_runtime().hipMemcpy(ctypes.c_void_p(self.grid_ptr(rel)), a.ctypes.data_as(ctypes.c_void_p),
                                      ctypes.c_size_t(a.nbytes), 1)  # hipMemcpyHostToDevice
        if rc != 0:
            raise PPPMError(f"hipMemcpy failed ({rc})")

    def get_history(self):
        """Host copies of the Neumann and Dirichlet rings, (T, 32) each."""
        pn, pd = ctypes.c_void_p(), ctypes.c_void_p()
        self._check(self.lib.pppm_history_ptrs(self.h, ctypes.byref(pn), ctypes.byref(pd)), "history_ptrs")
        out = []
        for p in (pn, pd):
            a = np.empty((self.nt, 32), dtype=np.float32)
            rc = _hip_runtime().hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(a.nbytes), 2)
            if rc != 0:
                raise PPPMError(f"hipMemcpy failed ({rc})")
            out.append(a)
        return out[0], out[1]

    def read_cells(self, xyz):
        xyz = np.ascontiguousarray(np.asarray(xyz, dtype=np.int32).reshape(-1, 3))
        out = np.zeros(xyz.shape[0], dtype=np.float32)
        self._check(self.lib.pppm_read_cells(self.h, xyz.ctypes.data_as(ctypes.c_void_p), xyz.shape[0],
                                             out.ctypes.data_as(ctypes.c_void_p)), "read_cells")
        return out

    def ffat_collect(self, w, fmap):
        """FFAT map accumulation at the current time level (collect_ffat_map,
        experiments/static/main.cu:60-80): fmap is a float32 device tensor of shape (6, 2w, 2w),
        updated in place on the solver's stream."""
        import torch
        if not (isinstance(fmap, torch.Tensor) and fmap.is_cuda and fmap.dtype == torch.float32
                and fmap.is_contiguous() and tuple(fmap.shape) == (6, 2 * w, 2 * w)):
            raise ValueError("fmap must be a contiguous float32 device tensor of shape (6, 2w, 2w)")
        self._check(self.lib.pppm_ffat_collect(self.h, int(w), ctypes.c_void_p(fmap.data_ptr())), "ffat_collect")

    def slice_image(self, axis, index, max_value=1.0):
        """save_grid's picture of plane `index` normal to `axis` (0: x, 1: y, 2: z) at the
        current level: (W, W, 4) uint8 RGBA in PNG row order (src/utils/visualize.h:120-132)."""
        w = ctypes.c_int()
        self._check(self.lib.pppm_slice_image(self.h, int(axis), int(index), float(max_value), None,
                                              ctypes.byref(w)), "slice_image")
        out = np.zeros((w.value, w.value, 4), np.uint8)
        self._check(self.lib.pppm_slice_image(self.h, int(axis), int(index), float(max_value),
                                              out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(w)), "slice_image")
        return out

    def save_grid(self, filename, max_value=1.0, slice_=(0.0, 0.5, 0.0)):
        """save_grid (src/utils/visualize.h:120-132): the plane int(N * slice_[a]) of the first
        axis a with a non-zero slice component, written as a PNG file."""
        a = next(i for i in range(3) if slice_[i] != 0)
        idx = int(np.float32(self._res) * np.float32(slice_[a]))
        write_png(filename, self.slice_image(a, idx, max_value))

    def get_triangles(self):
        out = np.zeros((self.nt, 13), dtype=np.float32)
        self._check(self.lib.pppm_triangles(self.h, out.ctypes.data_as(ctypes.c_void_p)), "triangles")
        return out

    def stats(self):
        out = np.zeros(8, dtype=np.int64)
        self._check(self.lib.pppm_stats(self.h, out.ctypes.data_as(ctypes.c_void_p)), "stats")
        keys = ["T", "C3", "P3", "B4", "P4", "grid_recomputes", "face_recomputes", "pool"]
        return dict(zip(keys, (int(x) for x in out)))

    def get_grid_weights(self):
        rec = np.zeros(self.nt * 27 * 2, dtype=LWH_DTYPE)
        self._check(self.lib.pppm_get_grid_weights(self.h, rec.ctypes.data_as(ctypes.c_void_p)), "get_grid_weights")
        return rec

    def set_grid_weights(self, rec):
        rec = np.ascontiguousarray(rec, dtype=LWH_DTYPE)
        self._check(self.lib.pppm_set_grid_weights(self.h, rec.ctypes.data_as(ctypes.c_void_p)), "set_grid_weights")

    def get_face_weights(self):
        st = self.stats()
        start = np.zeros(self.nt + 1, dtype=np.int32)
        nbr = np.zeros(max(st["P4"], 1), dtype=np.int32)
        rec = np.zeros(max(st["P4"], 1), dtype=LWH_DTYPE)
        self._check(self.lib.pppm_get_face_weights(self.h, start.ctypes.data_as(ctypes.c_void_p),
                                                   nbr.ctypes.data_as(ctypes.c_void_p),
                                                   rec.ctypes.data_as(ctypes.c_void_p)), "get_face_weights")
        return start, nbr[:st["P4"]], rec[:st["P4"]]

    def set_face_weights(self, rec):
        rec = np.ascontiguousarray(rec, dtype=LWH_DTYPE)
        self._check(self.lib.pppm_set_face_weights(self.h, rec.ctypes.data_as(ctypes.c_void_p)), "set_face_weights")

    def get_interp_weights(self):
        out = np.zeros((self.nt, 8), dtype=np.float32)
        self._check(self.lib.pppm_get_interp_weights(self.h, out.ctypes.data_as(ctypes.c_void_p)), "interp")
        return out

    def get_c3(self):
        st = self.stats()
        cells = np.zeros(max(st["C3"], 1), dtype=np.int32)
        start = np.zeros(st["C3"] + 1, dtype=np.int32)
        pairs = np.zeros(max(st["P3"], 1), dtype=np.int32)
        self._check(self.lib.pppm_get_c3(self.h, cells.ctypes.data_as(ctypes.c_void_p),
                                         start.ctypes.data_as(ctypes.c_void_p),
                                         pairs.ctypes.data_as(ctypes.c_void_p)), "get_c3")
        return cells[:st["C3"]], start, pairs[:st["P3"]]

    def get_b4(self):
        st = self.stats()
        cells = np.zeros(max(st["B4"], 1), dtype=np.int32)
        start = np.zeros(st["B4"] + 1, dtype=np.int32)
        self._check(self.lib.pppm_get_b4(self.h, cells.ctypes.data_as(ctypes.c_void_p),
                                         start.ctypes.data_as(ctypes.c_void_p), None), "get_b4")
        tris = np.zeros(max(int(start[-1]), 1), dtype=np.int32)
        self._check(self.lib.pppm_get_b4(self.h, None, None, tris.ctypes.data_as(ctypes.c_void_p)), "get_b4")
        return cells[:st["B4"]], start, tris[:int(start[-1])]

    # ------------------------------------------------------------ classification / fresh cells
    # GhostCellSolver::precompute_cell_data / fill_in_fresh_cell / set_solid_cell_zero
    # (src/algorithm/ghost_cell.h:69-212)
    AIR, SOLID, GHOST, UNKNOWN = 0, 1, 2, 3

    def classify(self, thin_shell=False):
        g = ctypes.c_int()
        self._check(self.lib.pppm_classify(self.h, int(bool(thin_shell)), ctypes.byref(g)), "classify")
        return g.value

    def fill_fresh(self, acc_now, acc_old):
        """Returns the number of fresh cells (-1: one had no nearest triangle)."""
        a = _arg(acc_now, "acc_now", self.nt, "f32")
        b = _arg(acc_old, "acc_old", self.nt, "f32")
        if a.dev != b.dev:
            raise PPPMError("acc_now and acc_old must both be host or both device")
        n = ctypes.c_int()
        self._check(self.lib.pppm_fill_fresh(self.h, a.ptr, b.ptr, a.dev, ctypes.byref(n)), "fill_fresh")
        return n.value

    def set_solid_zero(self):
        self._check(self.lib.pppm_set_solid_zero(self.h), "set_solid_zero")

    # ghost-cell boundary solve (GhostCellSolver, src/algorithm/ghost_cell.h:72-193)
    def ghost_setup(self, cond_threshold=25.0):
        """Stencil weights and sparse rows for the current classification; returns the number of
        second-order rows."""
        c = ctypes.c_int()
        self._check(self.lib.pppm_ghost_setup(self.h, float(cond_threshold), ctypes.byref(c)), "ghost_setup")
        return c.value

    def ghost_set_neumann(self, neumann):
        n = _arg(neumann, "neumann", self.nt, "f32")
        self._check(self.lib.pppm_ghost_set_neumann(self.h, n.ptr, n.dev), "ghost_set_neumann")

    def ghost_solve(self):
        """Right-hand side, solve and ghost values at the current level; returns the iterations."""
        it = ctypes.c_int()
        self._check(self.lib.pppm_ghost_solve(self.h, ctypes.byref(it)), "ghost_solve")
        return it.value

    def ghost_step(self, neumann):
        """GhostCellSolver::update: FDTD step, Neumann data, ghost solve, SOLID cells to zero."""
        n = _arg(neumann, "neumann", self.nt, "f32")
        self._check(self.lib.pppm_ghost_step(self.h, n.ptr, n.dev), "ghost_step")

    def ghost_update_mesh(self, verts):
        """GhostCellSolver::update_mesh; returns the fresh-cell count."""
        v = _arg(verts, "verts", 3 * self.nv, "f32")
        n = ctypes.c_int()
        self._check(self.lib.pppm_ghost_update_mesh(self.h, v.ptr, v.dev, ctypes.byref(n)), "ghost_update_mesh")
        return n.value

    def ghost_get(self, n):
        """Host copies of the ghost data (n = the classification's ghost count)."""
        out = {"cells": np.zeros(n, np.int32), "order": np.zeros(n, np.int32),
               "weights": np.zeros((n, 8), np.float32), "cond": np.zeros(n, np.float32),
               "rhs": np.zeros(n, np.float32), "x": np.zeros(n, np.float32),
               "cols": np.zeros((n, 9), np.int32), "vals": np.zeros((n, 9), np.float32)}
        ptrs = [out[k].ctypes.data_as(ctypes.c_void_p) for k in ("cells", "order", "weights", "cond", "rhs", "x",
                                                                  "cols", "vals")]
        self._check(self.lib.pppm_ghost_get(self.h, *ptrs), "ghost_get")
        return out

    def cells(self):
        NN = self._res ** 3
        out = {"type": np.zeros(NN, np.int32), "ghost_idx": np.zeros(NN, np.int32),
               "nearest": np.zeros(NN, np.int32), "point": np.zeros((NN, 3), np.float32),
               "reflect": np.zeros((NN, 3), np.float32), "dist": np.zeros(NN, np.float32)}
        ptrs = [out[k].ctypes.data_as(ctypes.c_void_p) for k in ("type", "ghost_idx", "nearest", "point",
                                                                  "reflect", "dist")]
        self._check(self.lib.pppm_get_cells(self.h, *ptrs), "get_cells")
        return out


class PPPMBatch:
    """Independent scenes stepped together (pppm_batch_*): every kernel of the time step is one
    launch over all scenes. `scenes` are PPPMSolver handles with their meshes set; the batch
    keeps references to them."""

    SIZE_CLASSES = (0, 40, 56, 72, 88, 112, 160)
    VARIANTS = ("cell", "tile8x64", "tile16x32", "tile4x64", "tile8x32")

    def __init__(self, scenes):
        self.lib = load_library()
        self.scenes = list(scenes)
        arr = (ctypes.c_void_p * len(self.scenes))(*[s.h.value for s in self.scenes])
        self.h = ctypes.c_void_p()
        rc = self.lib.pppm_batch_create(ctypes.byref(self.h), arr, len(self.scenes))
        if rc != 0:
            raise PPPMError(f"pppm_batch_create failed ({rc})")

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.pppm_batch_last_error(self.h)
            raise PPPMError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.lib.pppm_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def run(self, neumann_series, steps, listeners, out_dev, use_graph=True, strides=None):
        """neumann_series: one device tensor per scene (steps x stride_i); listeners: one (x, y, z)
        per scene, or for every scene the same number L of them (a list of L cells per scene);
        out_dev: device tensor of steps x n_scenes (x L) floats, row k = the listeners after step k."""
        n = len(self.scenes)
        if len(neumann_series) != n or len(listeners) != n:
            raise PPPMError("one Neumann series and one listener set per scene")
        strides = [s.nt for s in self.scenes] if strides is None else [int(x) for x in strides]
        args = []
        for i, (sc, ns, st) in enumerate(zip(self.scenes, neumann_series, strides)):
            if st < sc.nt:
                raise PPPMError(f"scene {i}: stride < triangle count")
            args.append(_arg(ns, f"neumann_series[{i}]", (int(steps) - 1) * st + sc.nt if steps > 0 else 0,
                             "f32", "device"))
        lx = np.asarray(listeners, dtype=np.int32)
        if lx.ndim == 2:
            lx = lx[:, None, :]
        if lx.ndim != 3 or lx.shape[0] != n or lx.shape[2] != 3 or not 1 <= lx.shape[1] <= 64:
            raise PPPMError("listeners: (x, y, z) per scene, or the same number (1 to 64) of cells per scene")
        L = lx.shape[1]
        lx = np.ascontiguousarray(lx)
        for i, sc in enumerate(self.scenes):
            if lx[i].min() < 0 or lx[i].max() >= sc.res():
                raise PPPMError(f"scene {i}: listener outside the grid")
        out = _arg(out_dev, "out", int(steps) * n * L, "f32", "device")
        ptrs = (ctypes.c_void_p * n)(*[a.ptr.value for a in args])
        strd = (ctypes.c_longlong * n)(*strides)
        self._check(self.lib.pppm_batch_run(self.h, ptrs, strd, int(steps), int(L), lx.ctypes.data_as(ctypes.c_void_p),
                                            out.ptr, int(use_graph)), "batch_run")

    def adjoint_reset(self, t_last):
        """Zero adjoint state of every scene; the next reverse step is forward step t_last."""
        self._check(self.lib.pppm_batch_adjoint_reset(self.h, int(t_last)), "batch_adjoint_reset")

    def adjoint_run(self, gout_dev, steps, listeners, grad_devs, use_graph=True, grad_strides=None):
        """Reverses `steps` steps of every scene: gout_dev (device, steps x n_scenes x L, the
        layout of run's output, forward time order; L listeners per scene as in run) -> grad_devs
        (one device tensor per scene, steps x grad_stride_i: the gradient of each step's Neumann
        input)."""
        n = len(self.scenes)
        if len(grad_devs) != n or len(listeners) != n:
            raise PPPMError("one gradient and one listener set per scene")
        if steps < 0:
            raise PPPMError("adjoint_run: steps >= 0")
        lx = np.asarray(listeners, dtype=np.int32)
        if lx.ndim == 2:
            lx = lx[:, None, :]
        if lx.ndim != 3 or lx.shape[0] != n or lx.shape[2] != 3 or not 0 <= lx.shape[1] <= 64:
            raise PPPMError("listeners: (x, y, z) per scene, or the same number (0 to 64) of cells per scene")
        L = lx.shape[1]
        lx = np.ascontiguousarray(lx)
        for i, sc in enumerate(self.scenes):
            if L and (lx[i].min() < 0 or lx[i].max() >= sc.res()):
                raise PPPMError(f"scene {i}: listener outside the grid")
        strides = [sc.nt for sc in self.scenes] if grad_strides is None else [int(x) for x in grad_strides]
        gr = []
        for i, (sc, g, st) in enumerate(zip(self.scenes, grad_devs, strides)):
            if st < sc.nt:
                raise PPPMError(f"scene {i}: grad_stride < triangle count")
            gr.append(_arg(g, f"grad[{i}]", (int(steps) - 1) * st + sc.nt if steps > 0 else 0, "f32", "device"))
        go = _arg(gout_dev, "gout", int(steps) * n * L, "f32", "device") if L else None
        ptrs = (ctypes.c_void_p * n)(*[a.ptr.value for a in gr])
        strd = (ctypes.c_longlong * n)(*strides)
        self._check(self.lib.pppm_batch_adjoint_run(self.h, go.ptr if go else None, int(steps), int(L),
                                                    lx.ctypes.data_as(ctypes.c_void_p), ptrs, strd, int(use_graph)),
                    "batch_adjoint_run")

    def autotune(self, reps=20):
        """Picks the fastest FDTD interior variant per size class; returns {class_lo: {variant: ms}}."""
        ms = np.full(len(self.SIZE_CLASSES) * len(self.VARIANTS), np.nan, np.float32)
        self._check(self.lib.pppm_batch_autotune(self.h, int(reps), ms.ctypes.data_as(ctypes.c_void_p)), "autotune")
        ms = ms.reshape(len(self.SIZE_CLASSES), len(self.VARIANTS))
        return {lo: {v: float(ms[c, k]) for k, v in enumerate(self.VARIANTS) if np.isfinite(ms[c, k])}
                for c, lo in enumerate(self.SIZE_CLASSES) if np.isfinite(ms[c]).any()}

    PHASES = ("fdtd", "windows", "far", "dirichlet", "near")

    def set_timing(self, enable=True):
        self._check(self.lib.pppm_batch_set_timing(self.h, int(enable)), "batch_set_timing")

    def get_timing(self):
        """{phase: (total_ms, launches)} of eager batch steps since the last call (HIP events)."""
        ms = np.zeros(8, np.float64)
        n = np.zeros(8, np.int64)
        self._check(self.lib.pppm_batch_get_timing(self.h, ms.ctypes.data_as(ctypes.c_void_p),
                                                   n.ctypes.data_as(ctypes.c_void_p)), "batch_get_timing")
        return {name: (float(ms[i]), int(n[i])) for i, name in enumerate(self.PHASES)}

    def variants(self):
        out = np.zeros(len(self.scenes), np.int32)
        self._check(self.lib.pppm_batch_variants(self.h, out.ctypes.data_as(ctypes.c_void_p)), "variants")
        return out

    def set_variants(self, v):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.int32), (len(self.scenes),)))
        self._check(self.lib.pppm_batch_set_variants(self.h, a.ctypes.data_as(ctypes.c_void_p)), "set_variants")


def write_png(path, rgba):
    """RGBA8 image (H, W, 4), top row first, as a PNG file (write_to_png,
    src/utils/array_writer.h:68-92): 8-bit RGBA, no interlace, deflate via zlib."""
    import struct
    import zlib
    img = np.ascontiguousarray(rgba, np.uint8)
    h, w = img.shape[0], img.shape[1]
    raw = b"".join(b"\x00" + img[r].tobytes() for r in range(h))

    def chunk(tag, data):
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    with open(path, "wb") as fh:
        fh.write(b"\x89PNG\r\n\x1a\n")
        fh.write(chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)))
        fh.write(chunk(b"IDAT", zlib.compress(raw, 6)))
        fh.write(chunk(b"IEND", b""))


def read_png(path):
    """Inverse of write_png for 8-bit RGBA images with filter type 0 rows."""
    import struct
    import zlib
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w, h = 8, b"", 0, 0
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        tag = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        if tag == b"IHDR":
            w, h = struct.unpack(">II", body[:8])
        elif tag == b"IDAT":
            idat += body
        pos += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 4 * w)
    assert np.all(raw[:, 0] == 0)
    return raw[:, 1:].reshape(h, w, 4).copy()


_hip = None


def _hip_runtime():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemcpy.restype = ctypes.c_int
    return _hip


class ModalSource:
    """Modal sound source of a rigid body on the GPU (RigidBody, src/utils/RigidBody.h:53-166):
    eig (3 nv, M) eigenvector matrix, eigvals (M,), the surface triangles and unit normals, the
    material's Rayleigh damping (alpha, beta) and the audio sample rate."""

    def __init__(self, eig, eigvals, tris, normals, alpha, beta, sample_rate, stream=None):
        self.lib = load_library()
        e = np.ascontiguousarray(eig, np.float32)
        ev = np.ascontiguousarray(eigvals, np.float32)
        t = np.ascontiguousarray(tris, np.int32)
        n = np.ascontiguousarray(normals, np.float32)
        self.nt, self.nmodes = t.shape[0], e.shape[1]
        self.h = ctypes.c_void_p()
        st = ctypes.c_void_p(stream) if stream else None
        rc = self.lib.pppm_modal_create(ctypes.byref(self.h), e.ctypes.data_as(ctypes.c_void_p), e.shape[0] // 3,
                                        self.nmodes, ev.ctypes.data_as(ctypes.c_void_p),
                                        t.ctypes.data_as(ctypes.c_void_p), self.nt, n.ctypes.data_as(ctypes.c_void_p),
                                        float(alpha), float(beta), float(sample_rate), st)
        if rc != 0:
            raise PPPMError(f"pppm_modal_create failed ({rc})")

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.pppm_modal_destroy(self.h)
            self.h = None

    def run(self, forces, accs):
        """forces (S, M) host array or device tensor; accs (S, T) float32 device tensor."""
        ns = int(forces.shape[0])
        f = _arg(forces, "forces", ns * self.nmodes, "f32")
        a = _arg(accs, "accs", ns * self.nt, "f32", "device")
        rc = self.lib.pppm_modal_run(self.h, f.ptr, ns, a.ptr, f.dev)
        if rc != 0:
            raise PPPMError(f"pppm_modal_run failed ({rc})")

    def _get(self, fn, shape):
        out = np.zeros(shape, np.float32)
        if fn(self.h, out.ctypes.data_as(ctypes.c_void_p)) != 0:
            raise PPPMError("modal query failed")
        return out

    def state(self):
        return self._get(self.lib.pppm_modal_state, (self.nmodes, 3))

    def set_state(self, q3):
        q = np.ascontiguousarray(q3, np.float32)
        if self.lib.pppm_modal_set_state(self.h, q.ctypes.data_as(ctypes.c_void_p)) != 0:
            raise PPPMError("pppm_modal_set_state failed")

    def surface_matrix(self):
        return self._get(self.lib.pppm_modal_surface_matrix, (self.nt, self.nmodes))

    def coeffs(self):
        return self._get(self.lib.pppm_modal_coeffs, (self.nmodes, 3))

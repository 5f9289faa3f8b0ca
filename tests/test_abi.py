"""The C ABI library: it loads without a GPU and exports exactly what include/pppm_hip.h
declares; the Python binding names the same set. No compute call is made here."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pppm_hip.h")
LIB = os.path.join(ROOT, "pppm-tdbem_amd", "libpppm_hip.so")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pppm_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        subprocess.run(["make", "-C", os.path.join(ROOT, "pppm-tdbem_amd"), "-s", "-j8"], check=True)
    return ctypes.CDLL(LIB)


def test_header_declares_the_boundary():
    names = header_functions()
    for required in ("pppm_create", "pppm_destroy", "pppm_set_mesh", "pppm_update_mesh", "pppm_step",
                     "pppm_fdtd_step", "pppm_solve_far", "pppm_set_neumann", "pppm_update_dirichlet",
                     "pppm_solve_near", "pppm_read_cells", "pppm_grid_ptr", "pppm_history_ptrs",
                     "pppm_triangles", "pppm_set_stream", "pppm_last_error"):
        assert required in names, required


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_the_header():
    import pppm_amd
    assert sorted(pppm_amd.EXPORTS) == header_functions()


def test_exports_are_unmangled_c_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    pppm = {n for n in exported if n.startswith("pppm_")}
    assert pppm == set(header_functions())
    assert not any(n.startswith("_Z") and "pppm" in n for n in exported)


def test_create_rejects_bad_arguments_without_touching_the_device(lib):
    lib.pppm_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_uint]
    lib.pppm_create.restype = ctypes.c_int
    h = ctypes.c_void_p()
    # argument checks come before any HIP call: res < 8, dl <= 0, 2*pml >= res
    assert lib.pppm_create(ctypes.byref(h), 4, 0.01, 1e-5, None, 0, 0.83, 343.0, 0) == -1
    assert lib.pppm_create(ctypes.byref(h), 32, 0.0, 1e-5, None, 0, 0.83, 343.0, 0) == -1
    assert lib.pppm_create(ctypes.byref(h), 32, 0.01, 1e-5, None, 16, 0.83, 343.0, 0) == -1
    assert lib.pppm_create(None, 32, 0.01, 1e-5, None, 0, 0.83, 343.0, 0) == -1
    assert not h.value

"""Argument checks of the Python binding (pppm_amd.py), on the CPU: a host array where device
memory is required, a short buffer, or a wrong element type must raise PPPMError before any C
entry point is called (a host pointer passed as device memory would make a kernel fault)."""
import numpy as np
import pytest


class _NoCalls:
    """Stands in for libpppm_hip.so: any C call is a test failure."""

    def __getattr__(self, name):
        raise AssertionError(f"{name} was called with an invalid argument")


def _solver(amd, nt=10, nv=8, res=16):
    s = amd.PPPMSolver.__new__(amd.PPPMSolver)
    s.lib, s.h, s.nt, s.nv, s._res = _NoCalls(), None, nt, nv, res
    return s


def test_arg_checks_type_size_and_memory(amd):
    import torch
    a = amd._arg(np.arange(6, dtype=np.float64), "x", 6, "f32")
    assert a.dev == 0 and a.keep.dtype == np.float32
    with pytest.raises(amd.PPPMError, match="needs 7"):
        amd._arg(np.zeros(6, np.float32), "x", 7)
    with pytest.raises(amd.PPPMError, match="integer"):
        amd._arg(np.zeros(3, np.float32), "tris", 3, "i32")
    with pytest.raises(amd.PPPMError, match="torch.float32"):
        amd._arg(torch.zeros(8, dtype=torch.float64), "x", 8)
    with pytest.raises(amd.PPPMError, match="contiguous"):
        amd._arg(torch.zeros(4, 4).t(), "x", 16)
    with pytest.raises(amd.PPPMError, match="device memory"):
        amd._arg(torch.zeros(8), "x", 8, "f32", "device")
    with pytest.raises(amd.PPPMError, match="device memory"):
        amd._arg(np.zeros(8, np.float32), "x", 8, "f32", "device")
    with pytest.raises(amd.PPPMError, match="required"):
        amd._arg(None, "x", 1)


def test_run_rejects_host_and_short_buffers(amd):
    import torch
    s = _solver(amd)
    host = np.zeros(5 * 10, np.float32)
    out = torch.zeros(5, 1)                       # CPU tensor: not device memory
    with pytest.raises(amd.PPPMError, match="out must be device memory"):
        s.run(host, 5, [(1, 1, 1)], out)
    with pytest.raises(amd.PPPMError, match="listener outside"):
        s.run(host, 5, [(1, 1, 16)], out)
    with pytest.raises(amd.PPPMError, match="stride"):
        s.run(host, 5, [], None, stride=9)
    # no listeners: the Neumann series itself must be device memory
    with pytest.raises(amd.PPPMError, match="neumann_series must be device memory"):
        s.run(host, 5, [], None)
    with pytest.raises(amd.PPPMError, match="verts_series"):
        s.run_moving(torch.zeros(5 * 24), torch.zeros(50), 5, [], None)


def test_step_and_update_inputs_are_checked(amd):
    import torch
    s = _solver(amd)
    with pytest.raises(amd.PPPMError, match="needs 10"):
        s.update_grid_and_face(np.zeros(9, np.float32))
    with pytest.raises(amd.PPPMError, match="needs 24"):
        s.update_mesh(np.zeros(23, np.float32))
    with pytest.raises(amd.PPPMError, match="both be host or both device"):
        s.update_mesh_and_step(np.zeros(24, np.float32), _FakeCuda(torch.zeros(10)))
    with pytest.raises(amd.PPPMError, match="acc_now"):
        s.fill_fresh(np.zeros(3, np.float32), np.zeros(10, np.float32))
    with pytest.raises(amd.PPPMError, match="needs 10"):
        s.set_neumann_condition(torch.zeros(4))


class _FakeCuda:
    """A CPU tensor that reports itself as device memory (for the host/device mix check only:
    its pointer never reaches the library because the check raises first)."""

    def __init__(self, t):
        self.t = t
        self.dtype = t.dtype
        self.is_cuda = True

    def data_ptr(self):
        return self.t.data_ptr()

    def is_contiguous(self):
        return True

    def numel(self):
        return self.t.numel()


def _batch(amd, scenes):
    b = amd.PPPMBatch.__new__(amd.PPPMBatch)
    b.lib, b.h, b.scenes = _NoCalls(), None, list(scenes)
    return b


def test_batch_run_and_adjoint_inputs_are_checked(amd):
    import torch
    sv = [_solver(amd, nt=10, res=16), _solver(amd, nt=12, res=20)]
    b = _batch(amd, sv)
    lis = [[(1, 1, 1)], [(2, 2, 2)]]
    host = [np.zeros(5 * 10, np.float32), np.zeros(5 * 12, np.float32)]
    out = torch.zeros(5, 2, 1)
    with pytest.raises(amd.PPPMError, match="one Neumann series and one listener set per scene"):
        b.run(host[:1], 5, lis, out)
    with pytest.raises(amd.PPPMError, match="must be device memory"):
        b.run(host, 5, lis, out)
    with pytest.raises(amd.PPPMError, match="stride < triangle count"):
        b.run(host, 5, lis, out, strides=[9, 12])
    grads = [torch.zeros(5, 10), torch.zeros(5, 12)]   # CPU tensors: not device memory
    gout = torch.zeros(5, 2, 1)
    with pytest.raises(amd.PPPMError, match="one gradient and one listener set per scene"):
        b.adjoint_run(gout, 5, lis, grads[:1])
    with pytest.raises(amd.PPPMError, match="steps >= 0"):
        b.adjoint_run(gout, -1, lis, grads)
    with pytest.raises(amd.PPPMError, match="listener outside the grid"):
        b.adjoint_run(gout, 5, [[(1, 1, 1)], [(2, 2, 20)]], grads)
    with pytest.raises(amd.PPPMError, match="grad_stride < triangle count"):
        b.adjoint_run(gout, 5, lis, grads, grad_strides=[9, 12])
    with pytest.raises(amd.PPPMError, match="must be device memory"):
        b.adjoint_run(gout, 5, lis, grads)
    with pytest.raises(amd.PPPMError, match="needs"):
        b.adjoint_run(gout, 5, lis, [torch.zeros(4, 10), torch.zeros(5, 12)])

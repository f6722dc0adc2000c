"""Argument validation of the Python bindings, without a GPU: a tensor or
array of the wrong dtype, layout or size is refused with ValueError before
its address reaches the library (a float32 or short array would otherwise
make the FP64 / int32 kernels read or write past the allocation)."""
import numpy as np
import pytest
import torch

from ucfd_sparse_amd import lib as L


def test_arr_accepts_matching_arrays():
    a = np.zeros(12)
    assert L.arr(a, np.float64, 12, "a") is a
    t = torch.zeros(12, dtype=torch.float64)
    assert L.arr(t, np.float64, 10, "t") is t
    i = L.arr([1, 2, 3], np.int32, 3, "i")
    assert i.dtype == np.int32 and i.flags.c_contiguous
    assert L.arr(None, np.float64, 5) is None
    assert L.arr(12345, np.float64, 5) == 12345          # raw address: caller's promise


@pytest.mark.parametrize("bad, msg", [
    (torch.zeros(12, dtype=torch.float32), "dtype"),
    (torch.zeros(12, dtype=torch.int64), "dtype"),
    (torch.zeros(11, dtype=torch.float64), "at least 12"),
    (torch.zeros(4, 6, dtype=torch.float64).t(), "contiguous"),
])
def test_arr_refuses_wrong_tensors(bad, msg):
    with pytest.raises(ValueError, match=msg):
        L.arr(bad, np.float64, 12, "x")


def test_arr_refuses_wrong_numpy_outputs_and_host_tensors_for_device_args():
    with pytest.raises(ValueError, match="output"):
        L.arr(np.zeros(12, np.float32), np.float64, 12, "y", out=True)
    with pytest.raises(ValueError, match="output"):
        L.arr(np.zeros((4, 6))[:, ::2], np.float64, 12, "y", out=True)
    with pytest.raises(ValueError, match="at least"):
        L.arr(np.zeros(3), np.float64, 12, "y")
    with pytest.raises(ValueError, match="CUDA"):
        L.arr(torch.zeros(12, dtype=torch.float64), np.float64, 12, "y", device=True)
    with pytest.raises(ValueError, match="CUDA"):
        L.arr(np.zeros(12), np.float64, 12, "y", device=True)


def test_solver_refuses_short_or_mistyped_mesh_arrays():
    from ucfd_sparse_amd.solver import LUSGSSolver
    N, F = 10, 4
    ok = dict(nei_ele=np.zeros(F * N, np.int32), fnorm_vol=np.zeros(F * N),
              vec_fnorm=np.zeros(3 * F * N), fspr=np.zeros(F * N), dt=np.zeros(N),
              upts=np.zeros(5 * N))
    for k, bad in (("vec_fnorm", np.zeros(F * N)), ("upts", np.zeros(4 * N)),
                   ("nei_ele", torch.zeros(F * N, dtype=torch.int64)),
                   ("dt", torch.zeros(N, dtype=torch.float32))):
        args = dict(ok, **{k: bad})
        with pytest.raises(ValueError, match=k):
            LUSGSSolver(N, F, [0, N], **args)


def test_krylov_handles_refuse_short_arrays():
    from ucfd_sparse_amd import krylov as K
    rowptr = np.array([0, 1, 2], np.int32)
    with pytest.raises(ValueError, match="colidx"):
        K.SpMat(rowptr, np.zeros(1, np.int32), np.zeros(50), 2, 5)
    with pytest.raises(ValueError, match="values"):
        K.SpMat(rowptr, np.zeros(2, np.int32), np.zeros(49), 2, 5)
    with pytest.raises(ValueError, match="diagslots"):
        K.Precon(rowptr, np.zeros(2, np.int32), np.zeros(1, np.int32), np.zeros(50), 2, 5)


def test_flowsys_refuses_short_structure():
    from ucfd_sparse_amd import flowsys as FS
    from ucfd_sparse_amd.mesh import FlowSysMesh
    nl, nf = 3, 4
    rowptr = np.array([0, 2, 4, 6], np.int32)
    fs = FlowSysMesh(nl, nf, rowptr, np.zeros(6, np.int32), np.zeros(6, np.int32),
                     np.ones(5, np.int8), np.ones(nf), np.ones(3 * nf), np.ones(nl),
                     np.zeros(2 * 25 * nf))
    with pytest.raises(ValueError, match="sides"):
        FS.FlowSys(fs, 5, 5)


def test_autograd_pass_refuses_host_and_mistyped_tensors():
    from ucfd_sparse_amd.autograd import ColoredSystem
    S = object.__new__(ColoredSystem)
    S.N, S.F = 8, 4
    dev_ok = {"rhsb": torch.zeros(40, dtype=torch.float64), "dt": torch.zeros(8,
              dtype=torch.float64), "fspr": torch.zeros(32, dtype=torch.float64)}
    with pytest.raises(ValueError, match="CUDA"):
        S.check_inputs(**dev_ok)
    with pytest.raises(ValueError, match="dtype"):
        S.check_inputs(torch.zeros(40, dtype=torch.float32), dev_ok["dt"], dev_ok["fspr"])

"""ctypes binding of the MI355X C-ABI library ``liblusgs.so``.

The argument types of the legacy LU-SGS entry points are exactly the ones the
reference's Python examples declare (``examples/omp3d.py:27-35``,
``examples/mpi3d.py:26-34``): ``c_int`` scalars, ``c_double`` factor and
``c_void_p`` arrays.  Arrays may be numpy arrays (host memory, staged through
the GPU by the library) or torch CUDA tensors / raw device addresses (used in
place).  There is no CPU fallback: loading fails loudly if the library is not
built, and every compute call fails with a nonzero status when no GPU is
present.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblusgs.so")
_LIB = None

c_int, c_double, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p

# name -> argtypes (restype c_int unless listed in _RESTYPE)
_SIGS = {
    # legacy array API, coloured (examples/omp3d.py:27-35, omp3d.c:180-189)
    "parallel_pre_lusgs": [c_int, c_int, c_double] + [c_void_p] * 4,
    "parallel_ns_lower_sweep": [c_int] * 4 + [c_void_p] * 9,
    "parallel_ns_upper_sweep": [c_int] * 4 + [c_void_p] * 9,
    "ns_parallel_lower_sweep": [c_int] * 4 + [c_void_p] * 10,
    "ns_parallel_upper_sweep": [c_int] * 4 + [c_void_p] * 10,
    "lusgs_parallel_ns_update": [c_int, c_void_p, c_void_p],
    "parallel_ns_lusgs_update": [c_int, c_void_p, c_void_p],
    # legacy array API, serial (examples/mpi3d.py:26-34, mpi3d.c:166-173)
    "serial_pre_lusgs": [c_int, c_int, c_double] + [c_void_p] * 4,
    "serial_ns_lower_sweep": [c_int, c_int] + [c_void_p] * 7,
    "serial_ns_upper_sweep": [c_int, c_int] + [c_void_p] * 7,
    "ns_serial_lower_sweep": [c_int, c_int] + [c_void_p] * 8,
    "ns_serial_upper_sweep": [c_int, c_int] + [c_void_p] * 8,
    "serial_ns_lusgs_update": [c_int, c_void_p, c_void_p],
    # RANS turbulence rows (doc/1_introduction.md:19-22; no reference binding)
    "parallel_rans_lower_sweep": [c_int] * 4 + [c_void_p] * 9,
    "parallel_rans_upper_sweep": [c_int] * 4 + [c_void_p] * 9,
    "rans_parallel_lower_sweep": [c_int] * 4 + [c_void_p] * 10,
    "rans_parallel_upper_sweep": [c_int] * 4 + [c_void_p] * 10,
    "serial_rans_lower_sweep": [c_int, c_int] + [c_void_p] * 7,
    "serial_rans_upper_sweep": [c_int, c_int] + [c_void_p] * 7,
    "rans_serial_lower_sweep": [c_int, c_int] + [c_void_p] * 8,
    "rans_serial_upper_sweep": [c_int, c_int] + [c_void_p] * 8,
    "lusgs_serial_ns_update": [c_int, c_void_p, c_void_p],
    # extensions (include/ucfd_ext.h)
    "ucfd_set_stream": [c_void_p],
    "ucfd_get_stream": [],
    "ucfd_synchronize": [],
    "ucfd_device_available": [],
    "ucfd_release_workspace": [],
    "ucfd_set_config": [c_int, c_int, c_int],
    "ucfd_get_config": [c_void_p, c_void_p, c_void_p],
    "ucfd_greedy_coloring": [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p],
    "ucfd_coloring_conflicts": [c_int, c_int, c_void_p, c_void_p],
    "ucfd_balanced_coloring": [c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p],
    "ucfd_greedy_coloring_ordered": [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p],
    "ucfd_bfs_order": [c_int, c_int, c_void_p, c_void_p, c_void_p],
    "ucfd_reduce_colors": [c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p],
    "ucfd_color_reorder": [c_int, c_void_p, c_void_p, c_void_p, c_void_p],
    "ucfd_permute_rows": [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p],
    "ucfd_permute_nei": [c_int, c_int, c_void_p, c_void_p, c_void_p],
    "ucfd_declare_colored_layout": [c_int, c_void_p, c_void_p, c_void_p],
    "ucfd_release_colored_layout": [c_void_p, c_void_p],
    "ucfd_lusgs_solver_create": [c_void_p, c_int, c_int, c_int] + [c_void_p] * 7,
    "ucfd_lusgs_solver_destroy": [c_void_p],
    "ucfd_lusgs_residual": [c_void_p] * 5,
    "ucfd_lusgs_residual_noalloc": [c_int, c_int, c_int] + [c_void_p] * 11,
    "ucfd_lusgs_solver_time_kernels": [c_void_p, c_void_p, c_int, c_void_p],
    "ucfd_lusgs_solve": [c_void_p, c_int, c_void_p, c_void_p, c_double, c_int, c_void_p],
    "ucfd_adjoint_create": [c_void_p, c_int, c_int, c_int] + [c_void_p] * 5,
    "ucfd_adjoint_destroy": [c_void_p],
    "ucfd_lusgs_adjoint": [c_void_p, c_double] + [c_void_p] * 8,
    # Krylov API (include/ucfdmat.h, ucfdprecon.h, ucfdsolver.h); handles are
    # passed by address (SpMat*, Precon*, Solver*) or by value (void*)
    "UCFDMatInit": [c_void_p],
    "UCFDMatDestroy": [c_void_p],
    "UCFDMatMult": [c_double, c_void_p, c_void_p, c_double, c_void_p],
    "UCFDMatCreateCSR": [c_void_p, c_int, c_void_p, c_void_p, c_void_p],
    "UCFDMatCreateBSR": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p],
    "UCFDMatCreateCUDACSR": [c_void_p, c_int, c_void_p, c_void_p, c_void_p],
    "UCFDMatCreateCUDABSR": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p],
    "UCFDPreconCreatefromArrays": [c_void_p] * 5,
    "UCFDPreconCreateNone": [c_void_p],
    "UCFDPreconPrepare": [c_void_p],
    "UCFDPreconDestroy": [c_void_p],
    "UCFDPreconSetBLUSGS": [c_void_p, c_int, c_int],
    "UCFDPreconSetPBLUSGS": [c_void_p, c_int, c_int, c_int, c_void_p],
    "UCFDPreconSetCUDABLUSGS": [c_void_p, c_int, c_int, c_int, c_void_p],
    "UCFDPreconSetILU": [c_void_p, c_int],
    "UCFDPreconSetBILU": [c_void_p, c_int, c_int],
    "UCFDPreconSetPBILU": [c_void_p, c_int, c_int, c_int, c_void_p],
    "UCFDPreconSetCUDABILU": [c_void_p, c_int, c_int, c_int, c_void_p],
    "UCFDPreconApply": [c_void_p, c_void_p],
    "UCFDPreconSetPrecision": [c_void_p, c_int],
    "UCFDSolverInit": [c_void_p],
    "UCFDSolverDestroy": [c_void_p],
    "UCFDSolve": [c_void_p] * 5,
    "UCFDSolverGetResult": [c_void_p] * 4,
    "UCFDSolverTraceResidualHistory": [c_void_p],
    "UCFDSolverGetResidualHistory": [c_void_p, c_void_p],
    "UCFDSolverCreateGMRES": [c_void_p, c_int, c_int, c_int, c_double],
    "UCFDSolverCreateCUDAGMRES": [c_void_p, c_int, c_int, c_int, c_double],
    "UCFDSolverCreateBICGSTAB": [c_void_p, c_int, c_int, c_double],
    "UCFDSolverSetRightPrecon": [c_void_p, ctypes.c_bool],
    # FlowSys API (include/ucfdlusgs.h)
    "UCFDFlowSysCreate": [c_void_p] + [c_int] * 6 + [c_void_p] * 7,
    "UCFDFlowSysSetElement": [c_void_p, c_int, c_int, c_int] + [c_void_p] * 7,
    "UCFDFlowSysDestroy": [c_void_p],
    "UCFDFlowSysSetLUSGS": [c_void_p, c_void_p, c_void_p],
    "UCFDLUSGS_Pack": [c_void_p, c_int, c_double, c_double],
    "UCFDLUSGS_Update": [c_void_p, c_int],
    "UCFDLUSGS_NSPrepare": [c_void_p, c_double],
    "UCFDLUSGS_RANSPrepare": [c_void_p, c_double],
    "UCFDLUSGS_NSLowerSweep": [c_void_p, c_double],
    "UCFDLUSGS_NSUpperSweep": [c_void_p, c_double],
    "UCFDLUSGS_RANSLowerSweep": [c_void_p, c_double],
    "UCFDLUSGS_RANSUpperSweep": [c_void_p, c_double],
    "UCFDLUSGS_Export": [c_void_p, c_int, c_void_p],
    "UCFDFlowSysSetBLUSGS": [c_void_p, c_void_p, c_void_p],
    "UCFDBLUSGS_Pack": [c_void_p, c_int, c_double],
    "UCFDBLUSGS_KWSST_Pack": [c_void_p, c_int, c_double, c_double],
    "UCFDBLUSGS_SA_Pack": [c_void_p, c_int, c_double, c_double],
    "UCFDBLUSGS_Update": [c_void_p, c_int],
    "UCFDBLUSGS_SubResidual": [c_void_p, c_int],
    "UCFDBLUSGS_NSPrepare": [c_void_p],
    "UCFDBLUSGS_RANSPrepare": [c_void_p],
    "UCFDBLUSGS_NSLowerSweep": [c_void_p],
    "UCFDBLUSGS_RANSLowerSweep": [c_void_p],
    "UCFDBLUSGS_NSUpperSweep": [c_void_p],
    "UCFDBLUSGS_RANSUpperSweep": [c_void_p],
    "UCFDBLUSGS_Reset": [c_void_p],
    "UCFDBLUSGS_SetPrecision": [c_void_p, c_int],
    "UCFDBLUSGS_Export": [c_void_p, c_int, c_void_p],
}

UCFD_SOLVE_FP64, UCFD_SOLVE_MIXED, UCFD_SOLVE_HIST = 0, 1, 128


class SolveStats(ctypes.Structure):
    """ucfd_solve_stats (include/ucfd_ext.h)."""
    _fields_ = [("converged", c_int), ("outer", c_int), ("passes", c_int), ("nhist", c_int),
                ("rnorm0", c_double), ("rel_residual", c_double),
                ("history", c_double * UCFD_SOLVE_HIST)]

    def hist(self):
        return [self.history[k] for k in range(self.nhist)]
_RESTYPE = {"ucfd_get_stream": c_void_p, "ucfd_coloring_conflicts": ctypes.c_longlong}


def _bind(lib, name, argtypes):
    fn = getattr(lib, name)
    fn.argtypes = argtypes
    fn.restype = _RESTYPE.get(name, c_int)


def extra_signatures():
    """Signatures registered by other modules of the package (block LU-SGS,
    Krylov, mixed precision, adjoint); they add to this table on import."""
    return _SIGS


def load():
    """Load liblusgs.so (once) and set every argtype; raises if missing.

    torch (when installed) is imported first: it ships its own HIP runtime
    (torch/lib/libamdhip64.so, soname libamdhip64.so.7).  Loaded first, it
    satisfies liblusgs.so's libamdhip64.so.7 dependency, so one process holds
    one HIP runtime and torch tensors and library calls share a device
    context.  Loading liblusgs.so first would pull in a second runtime from
    /opt/rocm and torch would then find no device."""
    global _LIB
    if _LIB is None:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: the HIP library is required (run "
                               "__graft_entry__.build() or make -C ucfd_sparse_amd/csrc)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, at in _SIGS.items():
            if hasattr(lib, name):
                _bind(lib, name, at)
        _LIB = lib
    return _LIB


def bind(name, argtypes, restype=c_int):
    """Register and bind one more entry point (used by sibling modules)."""
    _SIGS[name] = argtypes
    if restype is not c_int:
        _RESTYPE[name] = restype
    if _LIB is not None:
        _bind(_LIB, name, argtypes)


def device_count() -> int:
    return int(load().ucfd_device_available())


def ptr(a):
    """Raw address of a numpy array, torch tensor, ctypes object (or
    ctypes.byref of one) or integer address."""
    if a is None:
        return None
    if isinstance(a, int):
        return a
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    if isinstance(a, (ctypes._SimpleCData, ctypes.Structure, ctypes.Array)):
        return ctypes.addressof(a)
    if hasattr(a, "_obj"):                      # ctypes.byref(obj)
        return ctypes.addressof(a._obj)
    return a.ctypes.data


def arr(a, dtype, n=None, name="array", out=False, device=False):
    """Validate one array argument before its address reaches the library.

    dtype: the numpy dtype the entry point reads (float64, float32, int32,
    int8); n: the number of elements it touches (None: unchecked).  numpy
    inputs are converted to a C-contiguous array of dtype; numpy outputs
    (out=True) must already be one, since the results would otherwise land
    in a copy.  torch tensors must have exactly that dtype, be contiguous and
    hold at least n elements; device=True also requires a CUDA tensor.  Raw
    integer addresses pass unchecked.  Raises ValueError on a mismatch, so a
    wrong dtype or size never becomes an out-of-bounds device access."""
    if a is None or isinstance(a, int):
        return a
    want = np.dtype(dtype)
    if hasattr(a, "data_ptr"):
        import torch
        tdt = getattr(torch, want.name)
        if a.dtype != tdt:
            raise ValueError(f"{name}: dtype {a.dtype}, expected {tdt}")
        if not a.is_contiguous():
            raise ValueError(f"{name}: tensor must be contiguous")
        if device and not a.is_cuda:
            raise ValueError(f"{name}: must be a CUDA tensor")
        if n is not None and a.numel() < n:
            raise ValueError(f"{name}: {a.numel()} elements, at least {n} needed")
        return a
    if out:
        if not isinstance(a, np.ndarray) or a.dtype != want or not a.flags.c_contiguous:
            raise ValueError(f"{name}: output must be a C-contiguous {want} numpy array")
        b = a
    else:
        b = np.ascontiguousarray(a, want)
    if device:
        raise ValueError(f"{name}: must be a CUDA tensor")
    if n is not None and b.size < n:
        raise ValueError(f"{name}: {b.size} elements, at least {n} needed")
    return b


class UCFDError(RuntimeError):
    pass


def check(rc, what="ucfd call"):
    if rc != 0:
        raise UCFDError(f"{what} failed with status {rc}")
    return rc


def call(name, *args):
    """Call an entry point with numpy / torch / int arguments; raise on error."""
    fn = getattr(load(), name)
    conv = []
    for a, t in zip(args, fn.argtypes):
        conv.append(ptr(a) if t is c_void_p else a)
    rc = fn(*conv)
    if fn.restype is c_int:
        check(rc, name)
    return rc

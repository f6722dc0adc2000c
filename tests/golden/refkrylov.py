"""Drive the compiled reference Krylov library (oracle/_ref/libref_krylov.so,
built from src/krylov/*.c by `make -C oracle ref`) through its own public API:
UCFDMatCreateBSR, UCFDPreconCreatefromArrays + UCFDPreconSetBLUSGS /
UCFDPreconSetPBLUSGS, UCFDSolverCreateGMRES / UCFDSolverCreateBICGSTAB,
UCFDSolve, UCFDSolverGetResult, residual history.  OpenMP is pinned to one
thread so its reductions run in index order.  Only used where the reference
sources exist (fixture generation and oracle pins)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF = os.path.join(REPO, "oracle", "_ref", "libref_krylov.so")


def available():
    return os.path.exists(REF)


class _Ops(ctypes.Structure):
    _fields_ = [("prepare", ctypes.c_void_p), ("apply", ctypes.c_void_p),
                ("destroy", ctypes.c_void_p)]


class _Precon(ctypes.Structure):
    # src/krylov/include/impl/ucfdpcimpl.h:15-22
    _fields_ = [("type_name", ctypes.c_char_p), ("rowptr", ctypes.c_void_p),
                ("colidx", ctypes.c_void_p), ("diagslots", ctypes.c_void_p),
                ("values", ctypes.c_void_p), ("data", ctypes.c_void_p), ("ops", _Ops)]


_L = None


def single_thread_openmp():
    """Pin every OpenMP runtime loaded in this process to one thread.  The
    reference's parallel regions bind to whichever runtime exports the GOMP_*
    entry points first (libgomp, or LLVM's libomp when liblusgs.so was loaded
    before), so both are pinned; one thread makes its reductions run in index
    order, as the oracle's do."""
    libs = set()
    for line in open("/proc/self/maps"):
        path = line.split()[-1] if len(line.split()) >= 6 else ""
        base = os.path.basename(path)
        if base.startswith(("libgomp", "libomp", "libiomp")):
            libs.add(path)
    libs.add("libgomp.so.1")
    for path in libs:
        try:
            ctypes.CDLL(path).omp_set_num_threads(1)
        except OSError:
            pass


def load():
    global _L
    if _L is None:
        L = ctypes.CDLL(REF)
        single_thread_openmp()
        vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.UCFDMatCreateBSR.argtypes = [vp, ci, ci, vp, vp, vp]
        L.UCFDMatMult.argtypes = [cd, vp, vp, cd, vp]
        L.UCFDMatDestroy.argtypes = [vp]
        L.UCFDPreconCreatefromArrays.argtypes = [vp, vp, vp, vp, vp]
        L.UCFDPreconCreateNone.argtypes = [vp]
        L.UCFDPreconSetBLUSGS.argtypes = [vp, ci, ci]
        L.UCFDPreconSetPBLUSGS.argtypes = [vp, ci, ci, ci, vp]
        L.UCFDPreconSetILU.argtypes = [vp, ci]
        L.UCFDPreconSetBILU.argtypes = [vp, ci, ci]
        L.UCFDPreconSetPBILU.argtypes = [vp, ci, ci, ci, vp]
        L.UCFDPreconPrepare.argtypes = [vp]
        L.UCFDPreconDestroy.argtypes = [vp]
        L.UCFDSolverCreateGMRES.argtypes = [vp, ci, ci, ci, cd]
        L.UCFDSolverCreateBICGSTAB.argtypes = [vp, ci, ci, cd]
        L.UCFDSolve.argtypes = [vp, vp, vp, vp, vp]
        L.UCFDSolverGetResult.argtypes = [vp, vp, vp, vp]
        L.UCFDSolverTraceResidualHistory.argtypes = [vp]
        L.UCFDSolverGetResidualHistory.argtypes = [vp, vp]
        L.UCFDSolverDestroy.argtypes = [vp]
        L.ludcmp.argtypes = [ci, vp]
        L.lusub.argtypes = [ci, vp, vp]
        _L = L
    return _L


def spmv(A, x, alpha=1.0, beta=0.0, y=None):
    L = load()
    m = ctypes.c_void_p()
    L.UCFDMatCreateBSR(ctypes.byref(m), A.bn, A.blk, A.rowptr.ctypes.data, A.colidx.ctypes.data,
                       A.values.ctypes.data)
    y = np.zeros(A.bn * A.blk) if y is None else np.array(y, np.float64)
    xx = np.ascontiguousarray(x, np.float64)
    L.UCFDMatMult(alpha, m, xx.ctypes.data, beta, y.ctypes.data)
    L.UCFDMatDestroy(ctypes.byref(m))
    return y


def precon(A, icolors=None):
    """Returns (handle, keepalive); handle is a Precon prepared by the reference."""
    L = load()
    pc = ctypes.c_void_p()
    L.UCFDPreconCreatefromArrays(ctypes.byref(pc), A.rowptr.ctypes.data, A.colidx.ctypes.data,
                                 A.diagslots.ctypes.data, A.values.ctypes.data)
    keep = None
    if icolors is None:
        L.UCFDPreconSetBLUSGS(ctypes.byref(pc), A.bn, A.blk)
    else:
        keep = np.ascontiguousarray(icolors, np.int32)
        L.UCFDPreconSetPBLUSGS(ctypes.byref(pc), A.bn, A.blk, len(keep) - 1, keep.ctypes.data)
    L.UCFDPreconPrepare(pc)
    return pc, keep


def precon_apply(pc, b):
    """Calls the preconditioner's ops->apply (UCFDPreconApply is static inline
    in the reference, ucfdsolverimpl.h:63-67)."""
    P = ctypes.cast(pc, ctypes.POINTER(_Precon)).contents
    fn = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p)(P.ops.apply)
    x = np.ascontiguousarray(b, np.float64).copy()
    fn(pc, x.ctypes.data)
    return x


def ilu_precon(A, kind, icolors=None):
    """Returns (handle, keepalive): the reference's ILU-family preconditioner
    prepared (factored in place) on a copy of A's values."""
    L = load()
    vals = np.array(A.values, np.float64)
    pc = ctypes.c_void_p()
    L.UCFDPreconCreatefromArrays(ctypes.byref(pc), A.rowptr.ctypes.data, A.colidx.ctypes.data,
                                 A.diagslots.ctypes.data, vals.ctypes.data)
    keep = [vals]
    if kind == "ilu":
        L.UCFDPreconSetILU(ctypes.byref(pc), A.bn)
    elif kind == "bilu":
        L.UCFDPreconSetBILU(ctypes.byref(pc), A.bn, A.blk)
    else:
        ic = np.ascontiguousarray(icolors, np.int32)
        keep.append(ic)
        L.UCFDPreconSetPBILU(ctypes.byref(pc), A.bn, A.blk, len(ic) - 1, ic.ctypes.data)
    L.UCFDPreconPrepare(pc)
    return pc, keep


def solve(A, b, kind, m=30, maxiter=50, tol=1e-10, precon_kind="blusgs", icolors=None):
    """precon_kind: "none", "blusgs", or an ILU kind ("ilu", "bilu", "pbilu"
    with icolors)."""
    L = load()
    mat = ctypes.c_void_p()
    L.UCFDMatCreateBSR(ctypes.byref(mat), A.bn, A.blk, A.rowptr.ctypes.data,
                       A.colidx.ctypes.data, A.values.ctypes.data)
    keep = None
    if precon_kind == "none":
        pc = ctypes.c_void_p()
        L.UCFDPreconCreateNone(ctypes.byref(pc))
    elif precon_kind in ("ilu", "bilu", "pbilu"):
        pc, keep = ilu_precon(A, precon_kind, icolors)
    else:
        pc, _ = precon(A)
    s = ctypes.c_void_p()
    n = A.bn * A.blk
    if kind == "gmres":
        L.UCFDSolverCreateGMRES(ctypes.byref(s), n, m, maxiter, tol)
    else:
        L.UCFDSolverCreateBICGSTAB(ctypes.byref(s), n, maxiter, tol)
    L.UCFDSolverTraceResidualHistory(s)
    x = np.zeros(n)
    bb = np.ascontiguousarray(b, np.float64).copy()
    L.UCFDSolve(s, pc, mat, x.ctypes.data, bb.ctypes.data)
    st, it, res = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    L.UCFDSolverGetResult(s, ctypes.byref(st), ctypes.byref(it), ctypes.byref(res))
    hist = np.zeros(maxiter)
    L.UCFDSolverGetResidualHistory(s, hist.ctypes.data)
    L.UCFDSolverDestroy(ctypes.byref(s))
    if precon_kind != "none":
        L.UCFDPreconDestroy(ctypes.byref(pc))
    L.UCFDMatDestroy(ctypes.byref(mat))
    return x, it.value, st.value, res.value, hist


def ilu_prepare_apply(A, kind, b, icolors=None):
    """The reference's incomplete-LU preconditioners: kind "ilu" (scalar,
    A.blk == 1, UCFDPreconSetILU), "bilu" (UCFDPreconSetBILU) or "pbilu"
    (UCFDPreconSetPBILU with colour offsets icolors); UCFDPreconPrepare
    factors a copy of the values in place, then ops->apply runs on b.
    Returns (factored values, M^-1 b)."""
    L = load()
    vals = np.array(A.values, np.float64)
    pc = ctypes.c_void_p()
    L.UCFDPreconCreatefromArrays(ctypes.byref(pc), A.rowptr.ctypes.data, A.colidx.ctypes.data,
                                 A.diagslots.ctypes.data, vals.ctypes.data)
    keep = None
    if kind == "ilu":
        L.UCFDPreconSetILU(ctypes.byref(pc), A.bn)
    elif kind == "bilu":
        L.UCFDPreconSetBILU(ctypes.byref(pc), A.bn, A.blk)
    else:
        keep = np.ascontiguousarray(icolors, np.int32)
        L.UCFDPreconSetPBILU(ctypes.byref(pc), A.bn, A.blk, len(keep) - 1, keep.ctypes.data)
    L.UCFDPreconPrepare(pc)
    x = precon_apply(pc, b)
    L.UCFDPreconDestroy(ctypes.byref(pc))
    return vals, x

"""Python handles of the Krylov API of liblusgs.so (include/ucfdmat.h,
ucfdprecon.h, ucfdsolver.h): BSR / CSR matrices, block LU-SGS
preconditioners (sequential and coloured, FP64 or mixed FP32 apply), the
ILU(0) family (scalar, block, coloured block) and GMRES / BiCGSTAB.

The handles keep references to the arrays they were built from, because the
library keeps the pointers (as the reference does, src/krylov/matrix/
sparsemat.c:104-125, precon/precon.c:13-27).  Arrays may be numpy (host,
staged by the library on every call) or torch CUDA tensors (device, used in
place); vectors are FP64 of length bn*blk.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import lib as L

CONVERGED, HAPPYBREAKDOWN, REACH_ITERMAX, RHOBREAKDOWN, PIBREAKDOWN = 0, 1, 2, 3, 4
INITIALIZED = -1


def _i32(a, n=None, name="index array"):
    return L.arr(a, np.int32, n, name)


def _f64(a, n=None, name="value array"):
    return L.arr(a, np.float64, n, name)


def _nnz(rowptr, bn):
    """Stored blocks, rowptr[bn] (one device read for a tensor)."""
    return int(rowptr[bn])


class SpMat:
    """UCFDMatCreateBSR / UCFDMatCreateCSR (sparsemat.c:32-50, 104-125)."""

    def __init__(self, rowptr, colidx, values, bn, blk=None):
        self._lib = L.load()
        b2 = 1 if blk is None else int(blk) ** 2
        self.rowptr = _i32(rowptr, int(bn) + 1, "rowptr")
        nnz = _nnz(self.rowptr, int(bn))
        self.colidx = _i32(colidx, nnz, "colidx")
        self.values = _f64(values, nnz * b2, "values")
        h = ctypes.c_void_p()
        if blk is None:
            L.check(self._lib.UCFDMatCreateCSR(ctypes.byref(h), int(bn), L.ptr(self.rowptr),
                                               L.ptr(self.colidx), L.ptr(self.values)),
                    "UCFDMatCreateCSR")
            self.n = int(bn)
        else:
            L.check(self._lib.UCFDMatCreateBSR(ctypes.byref(h), int(bn), int(blk),
                                               L.ptr(self.rowptr), L.ptr(self.colidx),
                                               L.ptr(self.values)), "UCFDMatCreateBSR")
            self.n = int(bn) * int(blk)
        self.handle = h

    @classmethod
    def from_bsr(cls, A):
        """From ``ucfd_sparse_amd.mesh.BSR`` (or any object with bn, blk,
        rowptr, colidx, values)."""
        return cls(A.rowptr, A.colidx, A.values, A.bn, A.blk)

    def mult(self, x, y, alpha=1.0, beta=0.0):
        x = _f64(x, self.n, "x")
        y = L.arr(y, np.float64, self.n, "y", out=True)
        L.check(self._lib.UCFDMatMult(float(alpha), self.handle, L.ptr(x), float(beta), L.ptr(y)),
                "UCFDMatMult")
        return y

    def close(self):
        if self.handle:
            self._lib.UCFDMatDestroy(ctypes.byref(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Precon:
    """Block LU-SGS (``kind`` None or "blusgs") or ILU(0)-family ("ilu":
    scalar ILU(0) of a CSR matrix, blk 1, ilu.c:99-117; "bilu": block ILU(0),
    bilu.c:138-158) preconditioner.  ``icolors`` (colour offsets, length
    ncolors+1) selects the coloured form (pblusgs.c:107-127,
    pbilu.c:184-207), else the sequential one (blusgs.c:111-128); ``fp32``
    selects the mixed apply (block LU-SGS only).  The ILU kinds factor
    ``values`` in place at prepare, as the reference does: a float64
    contiguous array (or device tensor) passed in is overwritten."""

    def __init__(self, rowptr=None, colidx=None, diagslots=None, values=None, bn=0, blk=5,
                 icolors=None, fp32=False, kind=None):
        self._lib = L.load()
        h = ctypes.c_void_p()
        self.n = int(bn) * int(blk)
        if kind == "none":
            L.check(self._lib.UCFDPreconCreateNone(ctypes.byref(h)), "UCFDPreconCreateNone")
            self.handle = h
            return
        b2 = int(blk) ** 2
        self.rowptr = _i32(rowptr, int(bn) + 1, "rowptr")
        nnz = _nnz(self.rowptr, int(bn))
        self.colidx = _i32(colidx, nnz, "colidx")
        self.diagslots = _i32(diagslots, int(bn), "diagslots")
        self.values = _f64(values, nnz * b2, "values")
        L.check(self._lib.UCFDPreconCreatefromArrays(
            ctypes.byref(h), L.ptr(self.rowptr), L.ptr(self.colidx), L.ptr(self.diagslots),
            L.ptr(self.values)), "UCFDPreconCreatefromArrays")
        self.handle = h
        if kind == "ilu":
            if int(blk) != 1 or icolors is not None:
                raise ValueError("the scalar ILU(0) takes a CSR matrix (blk 1), no colours")
            L.check(self._lib.UCFDPreconSetILU(ctypes.byref(h), int(bn)), "UCFDPreconSetILU")
        elif kind == "bilu" and icolors is None:
            L.check(self._lib.UCFDPreconSetBILU(ctypes.byref(h), int(bn), int(blk)),
                    "UCFDPreconSetBILU")
        elif kind == "bilu":
            self.icolors = np.ascontiguousarray(icolors, np.int32)
            L.check(self._lib.UCFDPreconSetPBILU(ctypes.byref(h), int(bn), int(blk),
                                                 len(self.icolors) - 1, self.icolors.ctypes.data),
                    "UCFDPreconSetPBILU")
        elif kind not in (None, "blusgs"):
            raise ValueError(f"unknown preconditioner kind {kind!r}")
        elif icolors is None:
            L.check(self._lib.UCFDPreconSetBLUSGS(ctypes.byref(h), int(bn), int(blk)),
                    "UCFDPreconSetBLUSGS")
        else:
            self.icolors = np.ascontiguousarray(icolors, np.int32)
            L.check(self._lib.UCFDPreconSetPBLUSGS(ctypes.byref(h), int(bn), int(blk),
                                                   len(self.icolors) - 1, self.icolors.ctypes.data),
                    "UCFDPreconSetPBLUSGS")
        if fp32:
            L.check(self._lib.UCFDPreconSetPrecision(h, 1), "UCFDPreconSetPrecision")
        self.prepare()

    @classmethod
    def none(cls):
        return cls(kind="none")

    @classmethod
    def from_bsr(cls, A, icolors=None, fp32=False, kind=None):
        return cls(A.rowptr, A.colidx, A.diagslots, A.values, A.bn, A.blk, icolors, fp32, kind)

    def prepare(self):
        L.check(self._lib.UCFDPreconPrepare(self.handle), "UCFDPreconPrepare")

    def apply(self, b):
        b = L.arr(b, np.float64, self.n, "b", out=True)
        L.check(self._lib.UCFDPreconApply(self.handle, L.ptr(b)), "UCFDPreconApply")
        return b

    def close(self):
        if self.handle:
            self._lib.UCFDPreconDestroy(ctypes.byref(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Solver:
    """UCFDSolverCreateGMRES / UCFDSolverCreateBICGSTAB; ``right`` selects the
    right-preconditioned, relative-tolerance extension."""

    def __init__(self, kind, n, maxiter, tol, m=30, right=False, trace=True):
        self._lib = L.load()
        h = ctypes.c_void_p()
        if kind == "gmres":
            L.check(self._lib.UCFDSolverCreateGMRES(ctypes.byref(h), int(n), int(m), int(maxiter),
                                                    float(tol)), "UCFDSolverCreateGMRES")
        elif kind == "bicgstab":
            L.check(self._lib.UCFDSolverCreateBICGSTAB(ctypes.byref(h), int(n), int(maxiter),
                                                       float(tol)), "UCFDSolverCreateBICGSTAB")
        else:
            raise ValueError(kind)
        self.handle = h
        self.kind, self.n, self.maxiter = kind, int(n), int(maxiter)
        if right:
            L.check(self._lib.UCFDSolverSetRightPrecon(h, True), "UCFDSolverSetRightPrecon")
        if trace:
            L.check(self._lib.UCFDSolverTraceResidualHistory(h), "UCFDSolverTraceResidualHistory")
        self.trace = trace

    def solve(self, pc, A, x, b):
        x = L.arr(x, np.float64, self.n, "x", out=True)
        b = _f64(b, self.n, "b")
        L.check(self._lib.UCFDSolve(self.handle, pc.handle, A.handle, L.ptr(x), L.ptr(b)),
                "UCFDSolve")
        return self.result()

    def result(self):
        st, it, res = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        L.check(self._lib.UCFDSolverGetResult(self.handle, ctypes.byref(st), ctypes.byref(it),
                                              ctypes.byref(res)), "UCFDSolverGetResult")
        return st.value, it.value, res.value

    def history(self):
        h = np.zeros(max(self.maxiter, 1), np.float64)
        L.check(self._lib.UCFDSolverGetResidualHistory(self.handle, h.ctypes.data),
                "UCFDSolverGetResidualHistory")
        return h[:self.maxiter]

    def close(self):
        if self.handle:
            self._lib.UCFDSolverDestroy(ctypes.byref(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

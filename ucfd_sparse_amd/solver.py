"""Python handle of the defect-correction LU-SGS solver (FP64 and mixed
FP32/FP64 forms) of liblusgs.so (``ucfd_lusgs_solver_*``, include/ucfd_ext.h).

Arrays may be numpy (host; staged by the library) or torch CUDA tensors
(device; used in place), in the legacy layouts: ``b`` / ``dq`` are
``[5, neles]`` FP64.  The mesh must be colour-contiguous
(``ucfd_sparse_amd.layout.colored_layout``).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import lib as L


class LUSGSSolver:
    def __init__(self, neles, nface, color_off, nei_ele, fnorm_vol, vec_fnorm, fspr, dt, upts):
        self._lib = L.load()
        off = np.ascontiguousarray(color_off, np.int32)
        N, F = int(neles), int(nface)
        f64 = np.float64
        a = dict(nei=L.arr(nei_ele, np.int32, F * N, "nei_ele"),
                 fnv=L.arr(fnorm_vol, f64, F * N, "fnorm_vol"),
                 vfn=L.arr(vec_fnorm, f64, 3 * F * N, "vec_fnorm"),
                 fsp=L.arr(fspr, f64, F * N, "fspr"), dt=L.arr(dt, f64, N, "dt"),
                 upts=L.arr(upts, f64, 5 * N, "upts"))
        h = ctypes.c_void_p()
        L.check(self._lib.ucfd_lusgs_solver_create(
            ctypes.byref(h), N, F, len(off) - 1, off.ctypes.data,
            *[L.ptr(a[k]) for k in ("nei", "fnv", "vfn", "fsp", "dt", "upts")]),
            "ucfd_lusgs_solver_create")
        self.handle = h
        self.neles = N

    @classmethod
    def from_layout(cls, lay):
        m, f = lay.mesh, lay.flow
        return cls(m.neles, m.nface, lay.off, m.nei_ele, m.fnorm_vol, m.vec_fnorm, f.fspr, f.dt,
                   np.ascontiguousarray(f.upts[:5]))

    def residual(self, b, dq, r=None):
        n = 5 * self.neles
        b, dq = L.arr(b, np.float64, n, "b"), L.arr(dq, np.float64, n, "dq")
        r = L.arr(r, np.float64, n, "r", out=True)
        nrm = ctypes.c_double()
        L.check(self._lib.ucfd_lusgs_residual(self.handle, L.ptr(b), L.ptr(dq), L.ptr(r),
                                              ctypes.byref(nrm)), "ucfd_lusgs_residual")
        return nrm.value

    def solve(self, b, dq, mode="fp64", rtol=1e-10, maxiter=200):
        n = 5 * self.neles
        b, dq = L.arr(b, np.float64, n, "b"), L.arr(dq, np.float64, n, "dq", out=True)
        st = L.SolveStats()
        m = L.UCFD_SOLVE_FP64 if mode == "fp64" else L.UCFD_SOLVE_MIXED
        L.check(self._lib.ucfd_lusgs_solve(self.handle, m, L.ptr(b), L.ptr(dq), float(rtol),
                                           int(maxiter), ctypes.byref(st)), "ucfd_lusgs_solve")
        return st

    def close(self):
        if self.handle:
            self._lib.ucfd_lusgs_solver_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

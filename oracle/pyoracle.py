"""ORACLE — test infrastructure only.  ctypes binding of the CPU restatement
(oracle/liboracle.so, built by `make -C oracle`).  Imported only by tests/,
__graft_entry__.smoke() and the cpu_baseline leg of bench.py, as the checker.

Each function restates the reference function cited in the C source
(oracle/*_oracle.c); see there for the file:line it follows.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}


def lib(fast: bool = False, build: str | None = None):
    """The parity build (default), the CPU-baseline build (fast=True or
    build="fast") or the build with the reference's flags (build="ref")."""
    name = {None: "liboracle_fast.so" if fast else "liboracle.so", "parity": "liboracle.so",
            "fast": "liboracle_fast.so", "ref": "liboracle_ref.so"}[build]
    if name not in _LIBS:
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        L = ctypes.CDLL(path)
        vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.orc_flux.argtypes = [ci, ci, vp, vp, vp]
        L.orc_pre.argtypes = [ci, ci, cd, vp, vp, vp, vp]
        sweep = [ci, ci, ci, ci, ci, ci, ci, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.orc_lower.argtypes = sweep
        L.orc_upper.argtypes = sweep
        L.orc_update.argtypes = [ci, ci, vp, vp]
        L.orc_colored_iteration.argtypes = [ci, ci, ci, ci, ci, ci, vp, vp, vp, vp, vp, vp, vp,
                                            vp, vp, vp, vp]
        for fn in ("orc_pre", "orc_lower", "orc_upper", "orc_update", "orc_flux",
                   "orc_colored_iteration"):
            getattr(L, fn).restype = None
        L.orc_residual.argtypes = [ci, ci, ci] + [vp] * 9
        L.orc_residual.restype = cd
        L.orc_dc_solve.argtypes = [ci, ci, ci, ci] + [vp] * 11 + [cd, ci, vp]
        L.orc_dc_solve.restype = ci
        _LIBS[name] = L
    return _LIBS[name]


def _c(a, dtype):
    a = np.ascontiguousarray(a, dtype)
    return a


def flux(u, n):
    u = _c(u, np.float64)
    n = _c(n, np.float64)
    f = np.zeros_like(u)
    lib().orc_flux(len(u), len(n), u.ctypes.data, n.ctypes.data, f.ctypes.data)
    return f


def pre(fnorm_vol, dt, fspr, factor=1.0):
    fnv = _c(fnorm_vol, np.float64)
    F, N = fnv.shape
    diag = np.empty(N, np.float64)
    lib().orc_pre(N, F, factor, fnv.ctypes.data, _c(dt, np.float64).ctypes.data,
                  diag.ctypes.data, _c(fspr, np.float64).ctypes.data)
    return diag


def _sweep(fn, mode, n0, ne, nei, icolor, lcolor, fnorm_vol, vec_fnorm, upts, x_in, x_out,
           diag, fspr):
    F, N = nei.shape
    D = vec_fnorm.shape[1]
    args = [_c(a, np.int32) if a is not None else None for a in (nei, icolor, lcolor)]
    ptr = [a.ctypes.data if a is not None else None for a in args]
    fn(mode, n0, ne, N, F, D, D + 2, ptr[0], ptr[1], ptr[2],
       _c(fnorm_vol, np.float64).ctypes.data, _c(vec_fnorm, np.float64).ctypes.data,
       _c(upts, np.float64).ctypes.data, x_in.ctypes.data, x_out.ctypes.data,
       _c(diag, np.float64).ctypes.data, _c(fspr, np.float64).ctypes.data)


def lower(mode, n0, ne, nei, icolor, lcolor, fnorm_vol, vec_fnorm, upts, x_in, x_out, diag, fspr):
    """mode 0: coloured over icolor[n0:ne); mode 1: serial.  x_out is updated
    in place (x_in may be x_out)."""
    _sweep(lib().orc_lower, mode, n0, ne, nei, icolor, lcolor, fnorm_vol, vec_fnorm, upts, x_in,
           x_out, diag, fspr)


def upper(mode, n0, ne, nei, icolor, lcolor, fnorm_vol, vec_fnorm, upts, x_in, x_out, diag, fspr):
    _sweep(lib().orc_upper, mode, n0, ne, nei, icolor, lcolor, fnorm_vol, vec_fnorm, upts, x_in,
           x_out, diag, fspr)


def update(upts, x, nvars=None):
    nv = upts.shape[0] if nvars is None else nvars
    lib().orc_update(upts.shape[1], nv, upts.ctypes.data, _c(x, np.float64).ctypes.data)


def color_offsets(lcolor):
    counts = np.bincount(lcolor)[1:]
    return np.r_[0, np.cumsum(counts)].astype(np.int32)


def colored_pass(nei, icolor, lcolor, fnorm_vol, vec_fnorm, fspr, dt, upts, rhsb):
    """One pre/lower/upper/update pass of the coloured C in-place form.
    Returns (diag, dstar, dq, upts_new) without modifying the inputs."""
    F, N = nei.shape
    D = vec_fnorm.shape[1]
    nv = upts.shape[0]
    off = color_offsets(lcolor)
    diag = pre(fnorm_vol, dt, fspr)
    dub = np.ascontiguousarray(rhsb[:D + 2], np.float64).copy()
    for c in range(len(off) - 1):
        lower(0, off[c], off[c + 1], nei, icolor, lcolor, fnorm_vol, vec_fnorm, upts, dub, dub,
              diag, fspr)
    dstar = dub.copy()
    for c in reversed(range(len(off) - 1)):
        upper(0, off[c], off[c + 1], nei, icolor, lcolor, fnorm_vol, vec_fnorm, upts, dub, dub,
              diag, fspr)
    u = np.ascontiguousarray(upts, np.float64).copy()
    full = np.zeros((nv, N))
    full[:D + 2] = dub
    update(u, full)
    return diag, dstar, dub, u


def serial_pass(nei, fnorm_vol, vec_fnorm, fspr, dt, upts, rhsb):
    D = vec_fnorm.shape[1]
    diag = pre(fnorm_vol, dt, fspr)
    dub = np.ascontiguousarray(rhsb[:D + 2], np.float64).copy()
    lower(1, 0, 0, nei, None, None, fnorm_vol, vec_fnorm, upts, dub, dub, diag, fspr)
    dstar = dub.copy()
    upper(1, 0, 0, nei, None, None, fnorm_vol, vec_fnorm, upts, dub, dub, diag, fspr)
    u = np.ascontiguousarray(upts, np.float64).copy()
    full = np.zeros_like(u)
    full[:D + 2] = dub
    update(u, full)
    return diag, dstar, dub, u


def residual(nei, fnorm_vol, vec_fnorm, fspr, dt, q, b, dq):
    """r = b - A(dq) (oracle/residual_oracle.c); returns (r, ||r||)."""
    F, N = nei.shape
    D = vec_fnorm.shape[1]
    diag = pre(fnorm_vol, dt, fspr)
    r = np.empty((D + 2, N))
    c = lambda a: _c(a, np.float64)  # noqa: E731
    nrm = lib().orc_residual(N, F, D, _c(nei, np.int32).ctypes.data, c(fnorm_vol).ctypes.data,
                             c(vec_fnorm).ctypes.data, c(fspr).ctypes.data, diag.ctypes.data,
                             c(q).ctypes.data, c(b).ctypes.data, c(dq).ctypes.data, r.ctypes.data)
    return r, nrm


def dc_solve(nei, icolor, lcolor, fnorm_vol, vec_fnorm, fspr, dt, q, b, rtol, maxpass):
    """FP64 defect correction with coloured LU-SGS passes; returns
    (dq, passes, history)."""
    F, N = nei.shape
    D = vec_fnorm.shape[1]
    off = color_offsets(lcolor)
    dq = np.empty((D + 2, N))
    hist = np.zeros(maxpass + 1)
    c = lambda a: _c(a, np.float64)  # noqa: E731
    i32 = lambda a: _c(a, np.int32)  # noqa: E731
    n = lib().orc_dc_solve(N, F, D, len(off) - 1, off.ctypes.data, i32(nei).ctypes.data,
                           i32(icolor).ctypes.data, i32(lcolor).ctypes.data,
                           c(fnorm_vol).ctypes.data, c(vec_fnorm).ctypes.data, c(fspr).ctypes.data,
                           c(dt).ctypes.data, c(q[:D + 2]).ctypes.data, c(b[:D + 2]).ctypes.data,
                           dq.ctypes.data, rtol, maxpass, hist.ctypes.data)
    return dq, n, hist[:n + 1]


# ----------------------------------------------------------------- Krylov
def _klib():
    L = lib()
    if not hasattr(L, "_krylov_bound"):
        vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.orc_spmv_bsr.argtypes = [ci, ci, vp, vp, vp, cd, vp, cd, vp]
        L.orc_ludcmp.argtypes = [ci, vp]
        L.orc_lusub.argtypes = [ci, vp, vp]
        L.orc_blusgs_prepare.argtypes = [ci, ci, vp, vp, vp]
        L.orc_blusgs_apply.argtypes = [ci, ci, vp, vp, vp, vp, vp, ci, vp, vp]
        L.orc_gmres_solve.argtypes = [ci, ci] + [vp] * 5 + [ci, vp, ci, ci, cd, ci, vp, vp, vp,
                                                           vp, vp]
        L.orc_bicgstab_solve.argtypes = [ci, ci] + [vp] * 5 + [ci, vp, ci, cd, ci, vp, vp, vp,
                                                              vp, vp]
        for f in ("orc_spmv_bsr", "orc_ludcmp", "orc_lusub", "orc_blusgs_prepare",
                  "orc_blusgs_apply"):
            getattr(L, f).restype = None
        L.orc_gmres_solve_ilu.argtypes = [ci, ci] + [vp] * 4 + [ci, vp, ci, vp, ci, ci, cd, ci,
                                                               vp, vp, vp, vp, vp]
        L.orc_bicgstab_solve_ilu.argtypes = [ci, ci] + [vp] * 4 + [ci, vp, ci, vp, ci, cd, ci,
                                                                  vp, vp, vp, vp, vp]
        L.orc_gmres_solve.restype = ci
        L.orc_bicgstab_solve.restype = ci
        L.orc_gmres_solve_ilu.restype = ci
        L.orc_bicgstab_solve_ilu.restype = ci
        L._krylov_bound = True
    return L


def spmv_bsr(A, x, alpha=1.0, beta=0.0, y=None):
    y = np.zeros(A.bn * A.blk) if y is None else np.ascontiguousarray(y, np.float64).copy()
    _klib().orc_spmv_bsr(A.bn, A.blk, A.rowptr.ctypes.data, A.colidx.ctypes.data,
                         A.values.ctypes.data, alpha, _c(x, np.float64).ctypes.data, beta,
                         y.ctypes.data)
    return y


def ludcmp(Ab):
    a = np.ascontiguousarray(Ab, np.float64).copy()
    _klib().orc_ludcmp(a.shape[0], a.ctypes.data)
    return a


def lusub(LU, b):
    x = np.ascontiguousarray(b, np.float64).copy()
    _klib().orc_lusub(LU.shape[0], _c(LU, np.float64).ctypes.data, x.ctypes.data)
    return x


def blusgs_prepare(A):
    dv = np.empty((A.bn, A.blk, A.blk))
    _klib().orc_blusgs_prepare(A.bn, A.blk, A.diagslots.ctypes.data, A.values.ctypes.data,
                               dv.ctypes.data)
    return dv


def blusgs_apply(A, dv, b, icolors=None):
    x = np.ascontiguousarray(b, np.float64).copy()
    nc = 0 if icolors is None else len(icolors) - 1
    ic = None if icolors is None else _c(icolors, np.int32)
    _klib().orc_blusgs_apply(A.bn, A.blk, A.rowptr.ctypes.data, A.colidx.ctypes.data,
                             A.diagslots.ctypes.data, A.values.ctypes.data, dv.ctypes.data, nc,
                             None if ic is None else ic.ctypes.data, x.ctypes.data)
    return x


_ILU_KIND = {"bilu": 1, "pbilu": 1, "ilu": 2}


def _ilu_solve(fn, A, b, kind, icolors, *mid):
    """the oracle solvers with an ILU-family preconditioner factored from a
    copy of A's values"""
    f = ilu_prepare(A, kind)
    x = np.zeros(A.bn * A.blk)
    st, res = ctypes.c_int(), ctypes.c_double()
    maxiter = mid[-3]           # (m, maxiter, tol, right) or (maxiter, tol, relative)
    hist = np.zeros(maxiter + 1)
    nc = 0 if (icolors is None or kind != "pbilu") else len(icolors) - 1
    ic = None if nc == 0 else _c(icolors, np.int32)
    it = getattr(_klib(), fn)(A.bn, A.blk, A.rowptr.ctypes.data, A.colidx.ctypes.data,
                              A.diagslots.ctypes.data, A.values.ctypes.data, _ILU_KIND[kind],
                              f.ctypes.data, nc, None if ic is None else ic.ctypes.data, *mid,
                              x.ctypes.data, _c(b, np.float64).ctypes.data, ctypes.byref(st),
                              ctypes.byref(res), hist.ctypes.data)
    return x, it, st.value, res.value, hist[:it + 1]


def gmres_ilu(A, b, kind, m, maxiter, tol, right=False, icolors=None):
    """GMRES preconditioned by ILU(0) ("ilu"), block ILU(0) ("bilu") or its
    coloured form ("pbilu"); returns (x, restarts, stat, residual, history)"""
    return _ilu_solve("orc_gmres_solve_ilu", A, b, kind, icolors, m, maxiter, tol, int(right))


def bicgstab_ilu(A, b, kind, maxiter, tol, relative=False, icolors=None):
    return _ilu_solve("orc_bicgstab_solve_ilu", A, b, kind, icolors, maxiter, tol,
                      int(relative))


def gmres(A, b, m, maxiter, tol, right=False, precon=True, icolors=None, x0=None):
    """returns (x, restarts, stat, residual, history)"""
    dv = blusgs_prepare(A) if precon else None
    x = np.zeros(A.bn * A.blk) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
    st = ctypes.c_int()
    res = ctypes.c_double()
    hist = np.zeros(maxiter + 1)
    nc = 0 if icolors is None else len(icolors) - 1
    ic = None if icolors is None else _c(icolors, np.int32)
    it = _klib().orc_gmres_solve(A.bn, A.blk, A.rowptr.ctypes.data, A.colidx.ctypes.data,
                                 A.diagslots.ctypes.data, A.values.ctypes.data,
                                 None if dv is None else dv.ctypes.data, nc,
                                 None if ic is None else ic.ctypes.data, m, maxiter, tol,
                                 int(right), x.ctypes.data, _c(b, np.float64).ctypes.data,
                                 ctypes.byref(st), ctypes.byref(res), hist.ctypes.data)
    return x, it, st.value, res.value, hist[:it + 1]


def bicgstab(A, b, maxiter, tol, relative=False, precon=True, icolors=None):
    dv = blusgs_prepare(A) if precon else None
    x = np.zeros(A.bn * A.blk)
    st = ctypes.c_int()
    res = ctypes.c_double()
    hist = np.zeros(maxiter + 1)
    nc = 0 if icolors is None else len(icolors) - 1
    ic = None if icolors is None else _c(icolors, np.int32)
    it = _klib().orc_bicgstab_solve(A.bn, A.blk, A.rowptr.ctypes.data, A.colidx.ctypes.data,
                                    A.diagslots.ctypes.data, A.values.ctypes.data,
                                    None if dv is None else dv.ctypes.data, nc,
                                    None if ic is None else ic.ctypes.data, maxiter, tol,
                                    int(relative), x.ctypes.data, _c(b, np.float64).ctypes.data,
                                    ctypes.byref(st), ctypes.byref(res), hist.ctypes.data)
    return x, it, st.value, res.value, hist[:it + 1]


# ------------------------------------------------------ FlowSys block LU-SGS
def _fslib():
    L = lib()
    if not hasattr(L, "_fs_bound"):
        vp, ci, cd, cl = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int64
        L.orc_fs_pack.argtypes = [ci, ci, ci, cd, vp, vp, vp, vp, vp]
        L.orc_fs_tpack.argtypes = [ci, ci, ci, cd, cd, ci, vp, vp, vp, vp, vp]
        L.orc_fs_update.argtypes = [ci, ci, vp, vp, vp]
        L.orc_fs_subres.argtypes = [ci, ci, vp, vp, vp, vp, vp]
        L.orc_fs_prepare.argtypes = [ci, ci, cl, vp, vp, vp, vp, vp, vp, vp]
        L.orc_fs_sweep.argtypes = [ci, ci, ci, ci, ci, cl] + [vp] * 10
        for f in ("orc_fs_pack", "orc_fs_tpack", "orc_fs_update", "orc_fs_subres",
                  "orc_fs_prepare", "orc_fs_sweep"):
            getattr(L, f).restype = None
        L._fs_bound = True
    return L


class FlowSysOracle:
    """The FlowSys block LU-SGS state (rank rhs / du / dup / diag / tdiag) and
    its operations, restated in oracle/blusgs_oracle.c."""

    def __init__(self, fs, jmat, tjmat, el, nvars, nfvars):
        self.L = _fslib()
        self.fs, self.jmat, self.tjmat, self.el = fs, _c(jmat, np.float64), tjmat, el
        self.nvars, self.nfvars, self.nt = nvars, nfvars, nvars - nfvars
        N = fs.nlocal
        self.rhs = np.zeros(N * nvars)
        self.du = np.zeros(N * nvars)
        self.dup = np.zeros(N * nvars)
        self.diag = np.zeros(N * nfvars * nfvars)
        self.tdiag = np.zeros(N * self.nt * self.nt) if self.nt else None
        self.cid = _c(el["cell_ids"], np.int32)
        self.upts = np.array(el["uptsb"], np.float64)
        self.resid = np.zeros_like(self.upts)
        self.g = [_c(fs.rowptr, np.int32), _c(fs.colidx, np.int32), _c(fs.slots, np.int32),
                  _c(fs.sides, np.int8), _c(fs.face_area, np.float64), _c(fs.rcp_vol, np.float64)]

    def pack(self, a0):
        e = self.el
        self.L.orc_fs_pack(len(self.cid), self.nvars, self.nfvars, a0, self.cid.ctypes.data,
                           _c(e["rhs"], np.float64).ctypes.data, _c(e["dt"], np.float64).ctypes.data,
                           self.rhs.ctypes.data, self.diag.ctypes.data)

    def tpack(self, factor, a0, kind="sa"):
        e = self.el
        self.L.orc_fs_tpack(len(self.cid), self.nvars, self.nfvars, a0, factor,
                            1 if kind == "sa" else 2, self.cid.ctypes.data,
                            self.upts.ctypes.data, _c(e["dsrc"], np.float64).ctypes.data,
                            _c(e["dt"], np.float64).ctypes.data, self.tdiag.ctypes.data)

    def prepare(self, turb=False):
        rp, ci, sl, sd, fa, rv = self.g
        nv = self.nt if turb else self.nfvars
        jm = _c(self.tjmat if turb else self.jmat, np.float64)
        dg = self.tdiag if turb else self.diag
        self.L.orc_fs_prepare(self.fs.nlocal, nv, self.fs.nfaces, rp.ctypes.data, sl.ctypes.data,
                              sd.ctypes.data, fa.ctypes.data, rv.ctypes.data, dg.ctypes.data,
                              jm.ctypes.data)

    def sweep(self, upper, turb=False):
        rp, ci, sl, sd, fa, rv = self.g
        nv = self.nt if turb else self.nfvars
        var0 = self.nfvars if turb else 0
        jm = _c(self.tjmat if turb else self.jmat, np.float64)
        dg = self.tdiag if turb else self.diag
        self.L.orc_fs_sweep(int(upper), var0, nv, self.fs.nlocal, self.nvars, self.fs.nfaces,
                            rp.ctypes.data, ci.ctypes.data, sd.ctypes.data, sl.ctypes.data,
                            fa.ctypes.data, rv.ctypes.data, self.rhs.ctypes.data, dg.ctypes.data,
                            jm.ctypes.data, self.du.ctypes.data)

    def reset(self):
        self.du[:] = 0.0
        self.dup[:] = 0.0

    def subres(self):
        self.L.orc_fs_subres(len(self.cid), self.nvars, self.cid.ctypes.data,
                             _c(self.el["vol"], np.float64).ctypes.data, self.du.ctypes.data,
                             self.dup.ctypes.data, self.resid.ctypes.data)

    def update(self):
        self.L.orc_fs_update(len(self.cid), self.nvars, self.cid.ctypes.data, self.du.ctypes.data,
                             self.upts.ctypes.data)


def flowsys_run(fs, tjmat, el, nvars, nfvars, a0=0.3, turb_factor=1.5, kind="sa", iterations=2):
    """The oracle's version of tests/golden/refflowsys.run (same stages)."""
    O = FlowSysOracle(fs, fs.jmat, tjmat, el, nvars, nfvars)
    out = {}
    O.pack(a0)
    if O.nt:
        O.tpack(turb_factor, a0, kind)
        out["tdiag_pack"] = O.tdiag.copy()
    out["rank_rhs"] = O.rhs.copy()
    out["diag_pack"] = O.diag.copy()
    O.prepare()
    out["diag"] = O.diag.copy()
    if O.nt:
        O.prepare(turb=True)
        out["tdiag"] = O.tdiag.copy()
    O.reset()
    for it in range(iterations):
        O.sweep(False)
        out[f"du_lower{it}"] = O.du.copy()
        O.sweep(True)
        out[f"du_upper{it}"] = O.du.copy()
        if O.nt:
            O.sweep(False, turb=True)
            O.sweep(True, turb=True)
            out[f"du_rans{it}"] = O.du.copy()
        O.subres()
        out[f"resid{it}"] = O.resid.copy()
    O.update()
    out["upts"] = O.upts.copy()
    return out


# --------------------------------------------- incomplete-LU preconditioners
def _ilulib():
    L = lib()
    if not hasattr(L, "_ilu_bound"):
        vp, ci = ctypes.c_void_p, ctypes.c_int
        L.orc_ilu_prepare.argtypes = [ci, vp, vp, vp, vp]
        L.orc_ilu_apply.argtypes = [ci, vp, vp, vp, vp, vp]
        L.orc_bilu_prepare.argtypes = [ci, ci, vp, vp, vp, vp]
        L.orc_bilu_apply.argtypes = [ci, ci, vp, vp, vp, vp, vp]
        L.orc_pbilu_apply.argtypes = [ci, ci, ci, vp, vp, vp, vp, vp, vp]
        for f in ("orc_ilu_prepare", "orc_ilu_apply", "orc_bilu_prepare", "orc_bilu_apply",
                  "orc_pbilu_apply"):
            getattr(L, f).restype = None
        L._ilu_bound = True
    return L


def ilu_prepare(A, kind):
    """Factored copy of A.values: kind "ilu" (blk 1) or "bilu" / "pbilu"."""
    L = _ilulib()
    v = np.array(A.values, np.float64)
    rp, ci, ds = _c(A.rowptr, np.int32), _c(A.colidx, np.int32), _c(A.diagslots, np.int32)
    if kind == "ilu":
        L.orc_ilu_prepare(A.bn, rp.ctypes.data, ci.ctypes.data, ds.ctypes.data, v.ctypes.data)
    else:
        L.orc_bilu_prepare(A.bn, A.blk, rp.ctypes.data, ci.ctypes.data, ds.ctypes.data,
                           v.ctypes.data)
    return v


def ilu_apply(A, factored, b, kind, icolors=None):
    L = _ilulib()
    x = np.array(b, np.float64)
    f = _c(factored, np.float64)
    rp, ci, ds = _c(A.rowptr, np.int32), _c(A.colidx, np.int32), _c(A.diagslots, np.int32)
    if kind == "ilu":
        L.orc_ilu_apply(A.bn, rp.ctypes.data, ci.ctypes.data, ds.ctypes.data, f.ctypes.data,
                        x.ctypes.data)
    elif kind == "bilu":
        L.orc_bilu_apply(A.bn, A.blk, rp.ctypes.data, ci.ctypes.data, ds.ctypes.data,
                         f.ctypes.data, x.ctypes.data)
    else:
        ic = _c(icolors, np.int32)
        L.orc_pbilu_apply(A.bn, A.blk, len(ic) - 1, ic.ctypes.data, rp.ctypes.data,
                          ci.ctypes.data, ds.ctypes.data, f.ctypes.data, x.ctypes.data)
    return x


# ------------------------------------------------- FlowSys scalar LU-SGS
def _fsllib():
    L = lib()
    if not hasattr(L, "_fsl_bound"):
        vp, ci, cd, cl = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int64
        L.orc_fsl_pack.argtypes = [ci, ci, ci, ci, cd, cd] + [vp] * 8
        L.orc_fsl_update.argtypes = [ci, ci, ci, vp, vp, vp]
        L.orc_fsl_prepare.argtypes = [ci, ci, ci, cd] + [vp] * 6
        L.orc_fsl_sweep.argtypes = [ci, ci, cd, vp, ci, ci, ci, ci, cl] + [vp] * 10
        for f in ("orc_fsl_pack", "orc_fsl_update", "orc_fsl_prepare", "orc_fsl_sweep"):
            getattr(L, f).restype = None
        L._fsl_bound = True
    return L


class FlowSysLUSGSOracle:
    """The FlowSys scalar LU-SGS state (rank u / du / diag, [nvars, nlocal])
    and its operations, restated in oracle/flowsys_lusgs_oracle.c."""

    def __init__(self, fs, fspr, tfspr, el, nvars, nfvars):
        self.L = _fsllib()
        self.fs, self.el = fs, el
        self.fspr = _c(fspr, np.float64)
        self.tfspr = None if tfspr is None else _c(tfspr, np.float64)
        self.nvars, self.nfvars = nvars, nfvars
        N = fs.nlocal
        self.u = np.zeros(N * nvars)
        self.du = np.zeros(N * nvars)
        self.diag = np.zeros(N * nvars)
        self.cid = _c(el["cell_ids"], np.int32)
        self.upts = np.array(el["uptsb"], np.float64)
        self.g = [_c(fs.rowptr, np.int32), _c(fs.colidx, np.int32), _c(fs.sides, np.int8),
                  _c(fs.slots, np.int32), _c(fs.face_area, np.float64),
                  _c(fs.face_normal, np.float64), _c(fs.rcp_vol, np.float64)]

    def pack(self, turb_factor, a0):
        e = self.el
        self.L.orc_fsl_pack(self.fs.nlocal, len(self.cid), self.nvars, self.nfvars, turb_factor, a0,
                            self.cid.ctypes.data, self.upts.ctypes.data,
                            _c(e["rhs"], np.float64).ctypes.data, _c(e["dt"], np.float64).ctypes.data,
                            _c(e["dsrc"], np.float64).ctypes.data, self.u.ctypes.data,
                            self.du.ctypes.data, self.diag.ctypes.data)

    def prepare(self, kappa, turb=False):
        rp, ci, sd, sl, fa, fn, rv = self.g
        nv0, nv1 = (self.nfvars, self.nvars) if turb else (0, self.nfvars)
        lam = self.tfspr if turb else self.fspr
        self.L.orc_fsl_prepare(self.fs.nlocal, nv0, nv1, kappa, rp.ctypes.data, sl.ctypes.data,
                               fa.ctypes.data, rv.ctypes.data, lam.ctypes.data,
                               self.diag.ctypes.data)

    def sweep(self, upper, kappa, turb=False):
        rp, ci, sd, sl, fa, fn, rv = self.g
        lam = self.tfspr if turb else self.fspr
        self.L.orc_fsl_sweep(int(upper), int(turb), kappa, lam.ctypes.data, self.fs.nlocal,
                             self.nvars, self.nfvars, fn.shape[0], self.fs.nfaces, rp.ctypes.data,
                             ci.ctypes.data, sd.ctypes.data, sl.ctypes.data, fa.ctypes.data,
                             fn.ctypes.data, rv.ctypes.data, self.diag.ctypes.data,
                             self.u.ctypes.data, self.du.ctypes.data)

    def update(self):
        self.L.orc_fsl_update(self.fs.nlocal, len(self.cid), self.nvars, self.cid.ctypes.data,
                              self.du.ctypes.data, self.upts.ctypes.data)


def flowsys_lusgs_run(fs, fspr, tfspr, el, nvars, nfvars, a0=0.3, turb_factor=1.5, kappa=1.2,
                      iterations=2):
    """The stage sequence of tests/golden/refflowsys.run_lusgs on the oracle."""
    O = FlowSysLUSGSOracle(fs, fspr, tfspr, el, nvars, nfvars)
    nt = nvars - nfvars
    out = {}
    O.pack(turb_factor, a0)
    out["u"], out["du_pack"], out["diag_pack"] = O.u.copy(), O.du.copy(), O.diag.copy()
    O.prepare(kappa)
    out["diag_ns"] = O.diag.copy()
    if nt:
        O.prepare(kappa, turb=True)
        out["diag"] = O.diag.copy()
    for it in range(iterations):
        O.sweep(False, kappa)
        out[f"du_nslower{it}"] = O.du.copy()
        O.sweep(True, kappa)
        out[f"du_nsupper{it}"] = O.du.copy()
        if nt:
            O.sweep(False, kappa, turb=True)
            out[f"du_ranslower{it}"] = O.du.copy()
            O.sweep(True, kappa, turb=True)
            out[f"du_ransupper{it}"] = O.du.copy()
    O.update()
    out["upts"] = O.upts.copy()
    return out


# ---------------------------------------------------------------- adjoint
def _adjlib():
    L = lib()
    if not hasattr(L, "_adj_bound"):
        vp, ci, cd, cl = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int64
        L.orc_flux_jt.argtypes = [ci, vp, vp, vp, vp]
        L.orc_flux_jt.restype = None
        L.orc_lusgs_adjoint.argtypes = [cl, ci, ci, ci] + [vp] * 7 + [cd] + [vp] * 9
        L.orc_lusgs_adjoint.restype = None
        L._adj_bound = True
    return L


def flux_jt(u, n, v):
    """A(u; n)^T v (oracle/adjoint_oracle.c)."""
    out = np.zeros(len(u))
    _adjlib().orc_flux_jt(len(n), _c(u, np.float64).ctypes.data, _c(n, np.float64).ctypes.data,
                          _c(v, np.float64).ctypes.data, out.ctypes.data)
    return out


def face_pairs(nei):
    """pair[f, i] = face of nb = nei[f, i] pointing back to i (-1 on self faces)."""
    F, N = nei.shape
    pair = np.full((F, N), -1, np.int32)
    for f in range(F):
        nb = nei[f]
        inner = nb != np.arange(N)
        for g in range(F):
            hit = inner & (pair[f] < 0) & (nei[g, nb] == np.arange(N))
            pair[f, hit] = g
    return pair


def lusgs_adjoint(nei, off, fnorm_vol, vec_fnorm, fspr, dt, upts, dstar, dq, w, factor=1.0):
    """(rhsbar, dtbar, fsprbar) of J = sum(w * dq) for one coloured pass on a
    colour-contiguous numbering (colour c = cells off[c]..off[c+1]-1)."""
    F, N = nei.shape
    D = vec_fnorm.shape[1]
    nv = D + 2
    diag = pre(fnorm_vol, dt, fspr, factor)
    pair = face_pairs(nei)
    work = np.zeros((nv, N))
    rb = np.zeros((nv, N))
    dtb = np.zeros(N)
    fb = np.zeros((F, N))
    o = _c(off, np.int32)
    _adjlib().orc_lusgs_adjoint(
        N, F, D, len(o) - 1, o.ctypes.data, _c(nei, np.int32).ctypes.data, pair.ctypes.data,
        _c(fnorm_vol, np.float64).ctypes.data, _c(vec_fnorm, np.float64).ctypes.data,
        _c(fspr, np.float64).ctypes.data, _c(dt, np.float64).ctypes.data, factor,
        diag.ctypes.data, _c(upts[:nv], np.float64).ctypes.data,
        _c(dstar[:nv], np.float64).ctypes.data, _c(dq[:nv], np.float64).ctypes.data,
        _c(w[:nv], np.float64).ctypes.data, work.ctypes.data, rb.ctypes.data, dtb.ctypes.data,
        fb.ctypes.data)
    return rb, dtb, fb

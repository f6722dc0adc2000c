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


def lib(fast: bool = False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
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

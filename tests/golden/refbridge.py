"""Drive the compiled reference (oracle/_ref/libref_lusgs.so, built from the
reference's own src/lusgs sources) through the renumbering bridge.

The reference tree has no implementation of the legacy array API, only the
face-CSR FlowSys API with sequential sweeps (src/lusgs/lusgs.c:157-257).  A
coloured sweep (neighbour lower iff its colour is lower) equals that
sequential sweep run on the colour-ordered numbering, because colours are
contiguous in icolor, no two cells of a colour are adjacent and self faces
fail both tests.  So the legacy inputs are fed to UCFDFlowSysCreate as:
one CSR row per cell in rank order (rank = position in icolor), faces in
order f = 0..F-1, one slot per (cell, face) position, sides = +1,
face_area = fnorm_vol, rcp_vol = 1, face_normal = vec_fnorm, lambda = fspr,
kappa = turb_factor = 1, a0 = dsrc = 0.  With the identity rank the same
bridge gives the serial API.

Used only in the container that holds the reference (fixture generation and
the oracle-vs-reference tests); never on the GPU host.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF_LIB = os.path.join(REPO, "oracle", "_ref", "libref_lusgs.so")


class _FlowSys(ctypes.Structure):
    # src/lusgs/flowsys.h:34-59
    _fields_ = [("nlocal", ctypes.c_int), ("nvars", ctypes.c_int), ("nfvars", ctypes.c_int),
                ("nturbvars", ctypes.c_int), ("ndims", ctypes.c_int), ("nfaces", ctypes.c_int),
                ("rowptr", ctypes.c_void_p), ("colidx", ctypes.c_void_p),
                ("slots", ctypes.c_void_p), ("sides", ctypes.c_void_p),
                ("face_area", ctypes.c_void_p), ("face_normal", ctypes.c_void_p),
                ("rcp_vol", ctypes.c_void_p), ("data", ctypes.c_void_p),
                ("eles", ctypes.c_void_p), ("destroy", ctypes.c_void_p)]


class _LUSGSSys(ctypes.Structure):
    # src/lusgs/flowsys.h:20-24
    _fields_ = [("u", ctypes.POINTER(ctypes.c_double)), ("du", ctypes.POINTER(ctypes.c_double)),
                ("diag", ctypes.POINTER(ctypes.c_double)),
                ("fspr", ctypes.c_void_p), ("tfspr", ctypes.c_void_p)]


def available() -> bool:
    return os.path.exists(REF_LIB)


def load():
    lib = ctypes.CDLL(REF_LIB)
    vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.UCFDFlowSysCreate.argtypes = [vp, ci, ci, ci, ci, ci, ci, vp, vp, vp, vp, vp, vp, vp]
    lib.UCFDFlowSysSetLUSGS.argtypes = [vp, vp, vp]
    lib.UCFDFlowSysSetElement.argtypes = [vp, ci, ci, ci, vp, vp, vp, vp, vp, vp, vp]
    lib.UCFDFlowSysDestroy.argtypes = [vp]
    lib.UCFDLUSGS_Pack.argtypes = [vp, ci, cd, cd]
    lib.UCFDLUSGS_NSPrepare.argtypes = [vp, cd]
    lib.UCFDLUSGS_NSLowerSweep.argtypes = [vp, cd]
    lib.UCFDLUSGS_NSUpperSweep.argtypes = [vp, cd]
    lib.UCFDLUSGS_Update.argtypes = [vp, ci]
    lib.ns_flux_container.argtypes = [ci, ci, ci, vp, vp, vp]
    lib.ns_flux_container.restype = None
    return lib


def ref_flux(lib, u, n):
    """F(u; n) from the reference's ns_flux_container (flux.c:29-67)."""
    u = np.array(u, np.float64)
    n = np.ascontiguousarray(n, np.float64)
    f = np.zeros(len(u), np.float64)
    nd = len(n)
    lib.ns_flux_container(nd + 2, len(u) - nd - 2, nd, u.ctypes.data, n.ctypes.data, f.ctypes.data)
    return f


def lusgs_pass(lib, nei, fnorm_vol, vec_fnorm, fspr, dt, upts, rhs, rank=None, iterations=1,
               restore_rhs=True, keep=None):
    """Run `iterations` pre/lower/upper/update passes of the reference on the
    legacy inputs, with cells numbered by `rank` (rank[i] = row of cell i;
    identity = serial API).  Returns dict of legacy-layout arrays:
    diag [N], dstar [NV,N] (after lower), dq [NV,N] (after upper) of the
    first pass, and upts after every pass (list)."""
    F, N = nei.shape
    D = vec_fnorm.shape[1]
    NV = upts.shape[0]
    nfv = D + 2
    if rank is None:
        rank = np.arange(N, dtype=np.int32)
    rank = np.ascontiguousarray(rank, np.int32)
    inv = np.empty(N, np.int32)
    inv[rank] = np.arange(N, dtype=np.int32)          # inv[row] = cell

    nfaces = F * N
    rowptr = np.ascontiguousarray(np.arange(N + 1, dtype=np.int32) * F)
    # CSR row r = cell inv[r]; position r*F + f
    cells = inv
    colidx = np.ascontiguousarray(rank[nei[:, cells]].T.reshape(-1), np.int32)
    slots = np.arange(nfaces, dtype=np.int32)
    sides = np.ones(nfaces, np.int8)
    face_area = np.ascontiguousarray(fnorm_vol[:, cells].T.reshape(-1))
    face_normal = np.ascontiguousarray(
        np.stack([vec_fnorm[:, d, cells].T.reshape(-1) for d in range(D)]))
    rcp_vol = np.ones(N, np.float64)
    lam = np.ascontiguousarray(fspr[:, cells].T.reshape(-1))

    sysp = ctypes.c_void_p()
    lib.UCFDFlowSysCreate(ctypes.byref(sysp), N, NV, nfv, D, nfaces, 1, rowptr.ctypes.data,
                          colidx.ctypes.data, slots.ctypes.data, sides.ctypes.data,
                          face_area.ctypes.data, face_normal.ctypes.data, rcp_vol.ctypes.data)
    lib.UCFDFlowSysSetLUSGS(ctypes.byref(sysp), lam.ctypes.data, None)
    u = np.ascontiguousarray(upts.copy())
    rhs0 = np.ascontiguousarray(rhs.copy())
    rhs_cur = rhs0.copy()
    dsrc = np.zeros((NV, N), np.float64)
    vol = np.ones(N, np.float64)
    resid = np.zeros((NV, N), np.float64)
    dt = np.ascontiguousarray(dt, np.float64)
    cell_ids = rank.copy()
    lib.UCFDFlowSysSetElement(ctypes.byref(sysp), 0, N, F, cell_ids.ctypes.data, u.ctypes.data,
                              rhs_cur.ctypes.data, dt.ctypes.data, dsrc.ctypes.data,
                              vol.ctypes.data, resid.ctypes.data)
    fs = ctypes.cast(sysp, ctypes.POINTER(_FlowSys)).contents
    lsys = ctypes.cast(fs.data, ctypes.POINTER(_LUSGSSys)).contents

    def read(ptr):
        a = np.ctypeslib.as_array(ptr, shape=(NV * N,)).reshape(NV, N)
        return a[:nfv][:, rank].copy()       # back to legacy numbering

    out = {"upts": []}
    for it in range(iterations):
        if restore_rhs:
            rhs_cur[...] = rhs0
        lib.UCFDLUSGS_Pack(sysp, 0, 1.0, 0.0)
        lib.UCFDLUSGS_NSPrepare(sysp, 1.0)
        diag = np.ctypeslib.as_array(lsys.diag, shape=(NV * N,)).reshape(NV, N)[0][rank].copy()
        lib.UCFDLUSGS_NSLowerSweep(sysp, 1.0)
        dstar = read(lsys.du)
        lib.UCFDLUSGS_NSUpperSweep(sysp, 1.0)
        dq = read(lsys.du)
        lib.UCFDLUSGS_Update(sysp, 0)
        out["upts"].append(u.copy())
        if it == 0:
            out.update(diag=diag, dstar=dstar, dq=dq)
    lib.UCFDFlowSysDestroy(ctypes.byref(sysp))
    return out


def greedy_coloring_py(nei):
    """First-fit greedy colouring in ascending id (self faces ignored) —
    an independent Python statement of the build's colouring rule, for a
    symmetric neighbour table.  Returns (lcolor 1-based, icolor)."""
    F, N = nei.shape
    cols = [0] * N
    rows = [list(map(int, nei[f])) for f in range(F)]
    for i in range(N):
        used = set()
        for f in range(F):
            nb = rows[f][i]
            if nb < i:
                used.add(cols[nb])
        c = 1
        while c in used:
            c += 1
        cols[i] = c
    lcolor = np.array(cols, np.int32)
    icolor = np.argsort(lcolor, kind="stable").astype(np.int32)
    return lcolor, icolor

"""Drive the compiled reference's FlowSys block LU-SGS
(oracle/_ref/libref_lusgs.so, built from src/lusgs/*.c by `make -C oracle
ref`) through its own API and read its working arrays through the struct
layouts of src/lusgs/flowsys.h:27-59.  Used to write the FlowSys golden
fixture and, where the build exists, for live checks of the oracle."""
from __future__ import annotations

import ctypes

import numpy as np

import refbridge


class _BLUSGSSys(ctypes.Structure):
    # src/lusgs/flowsys.h:27-32
    _fields_ = [("rhs", ctypes.POINTER(ctypes.c_double)), ("du", ctypes.POINTER(ctypes.c_double)),
                ("dup", ctypes.POINTER(ctypes.c_double)), ("diag", ctypes.POINTER(ctypes.c_double)),
                ("tdiag", ctypes.POINTER(ctypes.c_double)), ("jmat", ctypes.c_void_p),
                ("tjmat", ctypes.c_void_p)]


def available() -> bool:
    return refbridge.available()


def load():
    lib = refbridge.load()
    vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.UCFDFlowSysSetBLUSGS.argtypes = [vp, vp, vp]
    lib.UCFDBLUSGS_Pack.argtypes = [vp, ci, cd]
    lib.UCFDBLUSGS_SA_Pack.argtypes = [vp, ci, cd, cd]
    lib.UCFDBLUSGS_KWSST_Pack.argtypes = [vp, ci, cd, cd]
    for f in ("UCFDBLUSGS_NSPrepare", "UCFDBLUSGS_RANSPrepare", "UCFDBLUSGS_NSLowerSweep",
              "UCFDBLUSGS_NSUpperSweep", "UCFDBLUSGS_RANSLowerSweep", "UCFDBLUSGS_RANSUpperSweep",
              "UCFDBLUSGS_Reset"):
        getattr(lib, f).argtypes = [vp]
    lib.UCFDBLUSGS_Update.argtypes = [vp, ci]
    lib.UCFDBLUSGS_SubResidual.argtypes = [vp, ci]
    return lib


def run(fs, tjmat, el, nvars, nfvars, a0=0.3, turb_factor=1.5, kind="sa", iterations=2):
    """One outer step of the documented sequence (Pack, turbulence Pack,
    NS/RANS Prepare, Reset, then per inner iteration NS lower/upper, RANS
    lower/upper, SubResidual; finally Update).  `el` holds the element
    arrays (cell_ids, uptsb, rhs, dt, dsrc, vol).  Returns the stages."""
    lib = load()
    N = fs.nlocal
    keep = []

    def c(a, dt):
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return a.ctypes.data
    sysp = ctypes.c_void_p()
    ndims = fs.face_normal.shape[0]
    lib.UCFDFlowSysCreate(ctypes.byref(sysp), N, nvars, nfvars, ndims, fs.nfaces, 1,
                          c(fs.rowptr, np.int32), c(fs.colidx, np.int32), c(fs.slots, np.int32),
                          c(fs.sides, np.int8), c(fs.face_area, np.float64),
                          c(fs.face_normal, np.float64), c(fs.rcp_vol, np.float64))
    lib.UCFDFlowSysSetBLUSGS(ctypes.byref(sysp), c(fs.jmat, np.float64),
                             None if tjmat is None else c(tjmat, np.float64))
    upts = np.array(el["uptsb"], np.float64)
    resid = np.zeros_like(upts)
    lib.UCFDFlowSysSetElement(ctypes.byref(sysp), 0, len(el["cell_ids"]), 4,
                              c(el["cell_ids"], np.int32), upts.ctypes.data,
                              c(el["rhs"], np.float64), c(el["dt"], np.float64),
                              c(el["dsrc"], np.float64), c(el["vol"], np.float64),
                              resid.ctypes.data)
    fsys = ctypes.cast(sysp, ctypes.POINTER(refbridge._FlowSys)).contents
    B = ctypes.cast(fsys.data, ctypes.POINTER(_BLUSGSSys)).contents
    nt = nvars - nfvars

    def arr(p, n):
        return np.ctypeslib.as_array(p, shape=(n,)).copy()
    out = {}
    lib.UCFDBLUSGS_Pack(sysp, 0, a0)
    if nt:
        (lib.UCFDBLUSGS_SA_Pack if kind == "sa" else lib.UCFDBLUSGS_KWSST_Pack)(
            sysp, 0, turb_factor, a0)
        out["tdiag_pack"] = arr(B.tdiag, N * nt * nt)
    out["rank_rhs"] = arr(B.rhs, N * nvars)
    out["diag_pack"] = arr(B.diag, N * nfvars * nfvars)
    lib.UCFDBLUSGS_NSPrepare(sysp)
    out["diag"] = arr(B.diag, N * nfvars * nfvars)
    if nt:
        lib.UCFDBLUSGS_RANSPrepare(sysp)
        out["tdiag"] = arr(B.tdiag, N * nt * nt)
    lib.UCFDBLUSGS_Reset(sysp)
    for it in range(iterations):
        lib.UCFDBLUSGS_NSLowerSweep(sysp)
        out[f"du_lower{it}"] = arr(B.du, N * nvars)
        lib.UCFDBLUSGS_NSUpperSweep(sysp)
        out[f"du_upper{it}"] = arr(B.du, N * nvars)
        if nt:
            lib.UCFDBLUSGS_RANSLowerSweep(sysp)
            lib.UCFDBLUSGS_RANSUpperSweep(sysp)
            out[f"du_rans{it}"] = arr(B.du, N * nvars)
        lib.UCFDBLUSGS_SubResidual(sysp, 0)
        out[f"resid{it}"] = resid.copy()
    lib.UCFDBLUSGS_Update(sysp, 0)
    out["upts"] = upts.copy()
    lib.UCFDFlowSysDestroy(ctypes.byref(sysp))
    return out


def run_lusgs(fs, fspr, tfspr, el, nvars, nfvars, a0=0.3, turb_factor=1.5, kappa=1.2,
              iterations=2):
    """The scalar LU-SGS sequence of the reference's FlowSys API
    (UCFDFlowSysSetLUSGS, UCFDLUSGS_Pack, NS / RANS Prepare, per iteration NS
    lower / upper and RANS lower / upper sweeps, Update; src/lusgs/lusgs.c),
    reading the rank arrays u / du / diag [nvars, nlocal] through the struct
    layout of src/lusgs/flowsys.h:20-24.  Returns the stages."""
    lib = load()
    vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.UCFDLUSGS_Update.argtypes = [vp, ci]
    for f in ("UCFDLUSGS_RANSPrepare", "UCFDLUSGS_NSLowerSweep", "UCFDLUSGS_NSUpperSweep",
              "UCFDLUSGS_RANSLowerSweep", "UCFDLUSGS_RANSUpperSweep"):
        getattr(lib, f).argtypes = [vp, cd]
    N = fs.nlocal
    keep = []

    def c(a, dt):
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return a.ctypes.data
    sysp = ctypes.c_void_p()
    ndims = fs.face_normal.shape[0]
    lib.UCFDFlowSysCreate(ctypes.byref(sysp), N, nvars, nfvars, ndims, fs.nfaces, 1,
                          c(fs.rowptr, np.int32), c(fs.colidx, np.int32), c(fs.slots, np.int32),
                          c(fs.sides, np.int8), c(fs.face_area, np.float64),
                          c(fs.face_normal, np.float64), c(fs.rcp_vol, np.float64))
    lib.UCFDFlowSysSetLUSGS(ctypes.byref(sysp), c(fspr, np.float64),
                            None if tfspr is None else c(tfspr, np.float64))
    upts = np.array(el["uptsb"], np.float64)
    resid = np.zeros_like(upts)
    lib.UCFDFlowSysSetElement(ctypes.byref(sysp), 0, len(el["cell_ids"]), 4,
                              c(el["cell_ids"], np.int32), upts.ctypes.data,
                              c(el["rhs"], np.float64), c(el["dt"], np.float64),
                              c(el["dsrc"], np.float64), c(el["vol"], np.float64),
                              resid.ctypes.data)
    fsys = ctypes.cast(sysp, ctypes.POINTER(refbridge._FlowSys)).contents
    Ls = ctypes.cast(fsys.data, ctypes.POINTER(refbridge._LUSGSSys)).contents
    nt = nvars - nfvars

    def arr(p):
        return np.ctypeslib.as_array(p, shape=(N * nvars,)).copy()
    out = {}
    lib.UCFDLUSGS_Pack(sysp, 0, turb_factor, a0)
    out["u"], out["du_pack"], out["diag_pack"] = arr(Ls.u), arr(Ls.du), arr(Ls.diag)
    lib.UCFDLUSGS_NSPrepare(sysp, kappa)
    out["diag_ns"] = arr(Ls.diag)
    if nt:
        lib.UCFDLUSGS_RANSPrepare(sysp, kappa)
        out["diag"] = arr(Ls.diag)
    for it in range(iterations):
        lib.UCFDLUSGS_NSLowerSweep(sysp, kappa)
        out[f"du_nslower{it}"] = arr(Ls.du)
        lib.UCFDLUSGS_NSUpperSweep(sysp, kappa)
        out[f"du_nsupper{it}"] = arr(Ls.du)
        if nt:
            lib.UCFDLUSGS_RANSLowerSweep(sysp, kappa)
            out[f"du_ranslower{it}"] = arr(Ls.du)
            lib.UCFDLUSGS_RANSUpperSweep(sysp, kappa)
            out[f"du_ransupper{it}"] = arr(Ls.du)
    lib.UCFDLUSGS_Update(sysp, 0)
    out["upts"] = upts.copy()
    lib.UCFDFlowSysDestroy(ctypes.byref(sysp))
    return out

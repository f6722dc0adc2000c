"""Python handle of the FlowSys block LU-SGS API of liblusgs.so
(include/ucfdlusgs.h; reference src/lusgs/flowsys.c, blusgs.c).

The handle keeps references to every array it was given, because the
library keeps the pointers (flowsys.c:18-54, 120-140).  Arrays may be numpy
(host; copied to the GPU for each call) or torch CUDA tensors (device; used
in place).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import lib as L

RHS, DU, DUP, DIAG, TDIAG = 0, 1, 2, 3, 4


def _npos(rowptr, nlocal):
    """Number of (cell, face) positions, rowptr[nlocal] (one device read for a
    tensor)."""
    return int(rowptr[nlocal])


def _structure(fs, ndims):
    """The face-CSR arrays of UCFDFlowSysCreate, validated (dtype, size)."""
    nl, nf = int(fs.nlocal), int(fs.nfaces)
    k = {"rowptr": L.arr(fs.rowptr, np.int32, nl + 1, "rowptr")}
    npos = _npos(k["rowptr"], nl)
    k["colidx"] = L.arr(fs.colidx, np.int32, npos, "colidx")
    k["slots"] = L.arr(fs.slots, np.int32, npos, "slots")
    k["sides"] = L.arr(fs.sides, np.int8, npos, "sides")
    k["face_area"] = L.arr(fs.face_area, np.float64, nf, "face_area")
    k["face_normal"] = L.arr(fs.face_normal, np.float64, int(ndims) * nf, "face_normal")
    k["rcp_vol"] = L.arr(fs.rcp_vol, np.float64, nl, "rcp_vol")
    return k


class FlowSys:
    def __init__(self, fs, nvars, nfvars, ndims=3, nelem=1, jmat=None, tjmat=None, fp32=False):
        """fs: ucfd_sparse_amd.mesh.FlowSysMesh (or any object with the same
        fields, numpy or torch)."""
        self._lib = L.load()
        self.keep = _structure(fs, ndims)
        k = self.keep
        self.nlocal, self.nvars, self.nfvars = int(fs.nlocal), int(nvars), int(nfvars)
        h = ctypes.c_void_p()
        L.check(self._lib.UCFDFlowSysCreate(
            ctypes.byref(h), self.nlocal, self.nvars, self.nfvars, int(ndims), int(fs.nfaces),
            int(nelem), L.ptr(k["rowptr"]), L.ptr(k["colidx"]), L.ptr(k["slots"]),
            L.ptr(k["sides"]), L.ptr(k["face_area"]), L.ptr(k["face_normal"]),
            L.ptr(k["rcp_vol"])), "UCFDFlowSysCreate")
        self.handle = h
        nf, nfv, ntv = int(fs.nfaces), self.nfvars, self.nvars - self.nfvars
        k["jmat"] = L.arr(fs.jmat if jmat is None else jmat, np.float64, 2 * nfv * nfv * nf,
                          "jmat")
        k["tjmat"] = None if tjmat is None else L.arr(tjmat, np.float64, 2 * ntv * ntv * nf,
                                                      "tjmat")
        L.check(self._lib.UCFDFlowSysSetBLUSGS(ctypes.byref(h), L.ptr(k["jmat"]),
                                               L.ptr(k["tjmat"])), "UCFDFlowSysSetBLUSGS")
        if fp32:
            L.check(self._lib.UCFDBLUSGS_SetPrecision(h, 1), "UCFDBLUSGS_SetPrecision")

    def set_element(self, eidx, el, nface=4):
        k = self.keep
        names = ("cell_ids", "uptsb", "rhs", "dt", "dsrc", "vol", "resid_out")
        ids = L.arr(el["cell_ids"], np.int32, None, "cell_ids")
        ne = len(ids)
        # rows of nvars element values ([nvars][neles]) or one value per element
        size = {"cell_ids": ne, "dt": ne, "vol": ne}
        for n in names:
            if n == "resid_out" and n not in el:
                k[(eidx, n)] = np.zeros_like(np.asarray(el["uptsb"])) \
                    if not hasattr(el["uptsb"], "data_ptr") else el["uptsb"].new_zeros(
                        el["uptsb"].shape)
            else:
                # numpy arrays are kept as converted copies: read written
                # ones (uptsb, resid_out) back with element_array()
                k[(eidx, n)] = L.arr(el[n], np.int32 if n == "cell_ids" else np.float64,
                                     size.get(n, self.nvars * ne), n)
        L.check(self._lib.UCFDFlowSysSetElement(
            ctypes.byref(self.handle), int(eidx), ne, int(nface),
            *[L.ptr(k[(eidx, n)]) for n in names]), "UCFDFlowSysSetElement")

    def element_array(self, eidx, name):
        return self.keep[(eidx, name)]

    def call(self, name, *args):
        L.check(getattr(self._lib, name)(self.handle, *args), name)

    def export(self, which, out=None):
        n = {RHS: self.nlocal * self.nvars, DU: self.nlocal * self.nvars,
             DUP: self.nlocal * self.nvars, DIAG: self.nlocal * self.nfvars ** 2,
             TDIAG: self.nlocal * (self.nvars - self.nfvars) ** 2}[which]
        out = np.zeros(n) if out is None else L.arr(out, np.float64, n, "out", out=True)
        L.check(self._lib.UCFDBLUSGS_Export(self.handle, int(which), L.ptr(out)),
                "UCFDBLUSGS_Export")
        return out

    def close(self):
        if self.handle:
            self._lib.UCFDFlowSysDestroy(ctypes.byref(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FlowSysLUSGS(FlowSys):
    """The scalar LU-SGS method of the FlowSys (UCFDFlowSysSetLUSGS):
    fspr / tfspr are per face slot (kept by the library).  Each Prepare also
    stores per-position face records that the sweeps after it read: the
    sweeps use the geometry and spectral radii of the last Prepare; sweeps
    with no Prepare before them read the arrays.  ndims as in
    UCFDFlowSysCreate (face_normal holds ndims * nfaces values)."""

    def __init__(self, fs, nvars, nfvars, fspr, tfspr=None, ndims=3, nelem=1):
        self._lib = L.load()
        self.keep = _structure(fs, ndims)
        k = self.keep
        self.nlocal, self.nvars, self.nfvars = int(fs.nlocal), int(nvars), int(nfvars)
        h = ctypes.c_void_p()
        L.check(self._lib.UCFDFlowSysCreate(
            ctypes.byref(h), self.nlocal, self.nvars, self.nfvars, int(ndims), int(fs.nfaces),
            int(nelem), L.ptr(k["rowptr"]), L.ptr(k["colidx"]), L.ptr(k["slots"]),
            L.ptr(k["sides"]), L.ptr(k["face_area"]), L.ptr(k["face_normal"]),
            L.ptr(k["rcp_vol"])), "UCFDFlowSysCreate")
        self.handle = h
        nf = int(fs.nfaces)
        k["fspr"] = L.arr(fspr, np.float64, nf, "fspr")
        k["tfspr"] = None if tfspr is None else L.arr(tfspr, np.float64, nf, "tfspr")
        L.check(self._lib.UCFDFlowSysSetLUSGS(ctypes.byref(h), L.ptr(k["fspr"]),
                                              L.ptr(k["tfspr"])), "UCFDFlowSysSetLUSGS")

    def export(self, which, out=None):
        n = self.nlocal * self.nvars
        out = np.zeros(n) if out is None else L.arr(out, np.float64, n, "out", out=True)
        L.check(self._lib.UCFDLUSGS_Export(self.handle, int(which), L.ptr(out)),
                "UCFDLUSGS_Export")
        return out


def run_lusgs(sysobj, a0=0.3, turb_factor=1.5, kappa=1.2, iterations=2):
    """The stage sequence of tests/golden/refflowsys.run_lusgs on a
    FlowSysLUSGS handle with element 0 set; returns host copies."""
    s = sysobj
    nt = s.nvars - s.nfvars
    out = {}
    s.call("UCFDLUSGS_Pack", 0, float(turb_factor), float(a0))
    out["u"], out["du_pack"], out["diag_pack"] = s.export(0), s.export(1), s.export(2)
    s.call("UCFDLUSGS_NSPrepare", float(kappa))
    out["diag_ns"] = s.export(2)
    if nt:
        s.call("UCFDLUSGS_RANSPrepare", float(kappa))
        out["diag"] = s.export(2)
    for it in range(iterations):
        s.call("UCFDLUSGS_NSLowerSweep", float(kappa))
        out[f"du_nslower{it}"] = s.export(1)
        s.call("UCFDLUSGS_NSUpperSweep", float(kappa))
        out[f"du_nsupper{it}"] = s.export(1)
        if nt:
            s.call("UCFDLUSGS_RANSLowerSweep", float(kappa))
            out[f"du_ranslower{it}"] = s.export(1)
            s.call("UCFDLUSGS_RANSUpperSweep", float(kappa))
            out[f"du_ransupper{it}"] = s.export(1)
    s.call("UCFDLUSGS_Update", 0)
    u = s.element_array(0, "uptsb")
    out["upts"] = (u.cpu().numpy() if hasattr(u, "data_ptr") else u).copy()
    return out


def run(sysobj, a0=0.3, turb_factor=1.5, kind="sa", iterations=2):
    """The stage sequence of tests/golden/refflowsys.run on a FlowSys handle
    with element 0 set; returns host copies of every stage."""
    s = sysobj
    nt = s.nvars - s.nfvars
    out = {}
    s.call("UCFDBLUSGS_Pack", 0, float(a0))
    if nt:
        s.call("UCFDBLUSGS_SA_Pack" if kind == "sa" else "UCFDBLUSGS_KWSST_Pack", 0,
               float(turb_factor), float(a0))
        out["tdiag_pack"] = s.export(TDIAG)
    out["rank_rhs"] = s.export(RHS)
    out["diag_pack"] = s.export(DIAG)
    s.call("UCFDBLUSGS_NSPrepare")
    out["diag"] = s.export(DIAG)
    if nt:
        s.call("UCFDBLUSGS_RANSPrepare")
        out["tdiag"] = s.export(TDIAG)
    s.call("UCFDBLUSGS_Reset")
    for it in range(iterations):
        s.call("UCFDBLUSGS_NSLowerSweep")
        out[f"du_lower{it}"] = s.export(DU)
        s.call("UCFDBLUSGS_NSUpperSweep")
        out[f"du_upper{it}"] = s.export(DU)
        if nt:
            s.call("UCFDBLUSGS_RANSLowerSweep")
            s.call("UCFDBLUSGS_RANSUpperSweep")
            out[f"du_rans{it}"] = s.export(DU)
        s.call("UCFDBLUSGS_SubResidual", 0)
        r = s.element_array(0, "resid_out")
        out[f"resid{it}"] = (r.cpu().numpy() if hasattr(r, "data_ptr") else r).reshape(-1).copy()
    s.call("UCFDBLUSGS_Update", 0)
    u = s.element_array(0, "uptsb")
    out["upts"] = (u.cpu().numpy() if hasattr(u, "data_ptr") else u).copy()
    return out

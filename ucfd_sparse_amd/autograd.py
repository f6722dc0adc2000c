"""torch binding of the differentiation form: one FP64 coloured LU-SGS pass
dQ = upper(lower(pre(dt, fspr), rhsb)) as a torch.autograd.Function whose
backward is the library's hand-written adjoint (ucfd_lusgs_adjoint,
include/ucfd_ext.h).

The forward runs the legacy API (parallel_pre_lusgs, ns_parallel_lower_sweep
per colour, ns_parallel_upper_sweep per colour in reverse) on device tensors
of a colour-contiguous layout, on torch's current stream.  The lower sweep's
output dQ* is either kept for the backward (two-array sweeps: rhsb -> dQ*,
dQ* -> dQ) or, with recompute=True, never kept: the forward runs the
in-place sweeps on one buffer (rhsb copy -> dQ* -> dQ) and the backward
recomputes dQ* (pre + lower sweeps again) and lets the adjoint overwrite it
with the rhsb gradient.  That saves 5 x neles doubles both between forward
and backward and at the peak, for half a forward of extra work.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import lib as L


class _Stream:
    """Route library calls to torch's current stream for the duration."""

    def __init__(self, lib):
        self.lib = lib

    def __enter__(self):
        self.prev = self.lib.ucfd_get_stream()
        self.lib.ucfd_set_stream(torch.cuda.current_stream().cuda_stream)

    def __exit__(self, *exc):
        self.lib.ucfd_set_stream(self.prev)
        return False


class ColoredSystem:
    """Device-resident mesh of a colour-contiguous layout
    (ucfd_sparse_amd.layout.ColoredLayout): nei_ele, fnorm_vol, vec_fnorm,
    upts, identity icolor / non-decreasing lcolor, declared to the library."""

    def __init__(self, lay, device="cuda:0"):
        self.lib = L.load()
        m, f = lay.mesh, lay.flow
        dev = torch.device(device)

        def up(a):
            return torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        self.N, self.F = m.neles, m.nface
        self.nei, self.fnv, self.vfn = up(m.nei_ele), up(m.fnorm_vol), up(m.vec_fnorm)
        self.upts = up(f.upts)
        self.icolor, self.lcolor = up(lay.icolor), up(lay.lcolor)
        self.off = [int(x) for x in lay.off]
        L.check(self.lib.ucfd_declare_colored_layout(self.N, self.icolor.data_ptr(),
                                                     self.lcolor.data_ptr(), None),
                "ucfd_declare_colored_layout")
        h = ctypes.c_void_p()
        off = np.ascontiguousarray(lay.off, np.int32)
        L.check(self.lib.ucfd_adjoint_create(ctypes.byref(h), self.N, self.F, len(off) - 1,
                                             off.ctypes.data, self.nei.data_ptr(),
                                             self.fnv.data_ptr(), self.vfn.data_ptr(),
                                             self.upts.data_ptr()), "ucfd_adjoint_create")
        self.handle = h

    def check_inputs(self, rhsb, dt, fspr, **outs):
        """Every array the pass reads or writes: float64, contiguous, on the
        GPU, at least as large as the kernels index (ValueError otherwise)."""
        N, F = self.N, self.F
        if rhsb is not None:
            L.arr(rhsb, np.float64, 5 * N, "rhsb", device=True)
        L.arr(dt, np.float64, N, "dt", device=True)
        L.arr(fspr, np.float64, F * N, "fspr", device=True)
        sizes = {"dq": 5 * N, "dstar": 5 * N, "diag": N, "w": 5 * N}
        for k, v in outs.items():
            if v is not None:
                L.arr(v, np.float64, sizes[k], k, device=True)

    def forward_arrays(self, rhsb, dt, fspr, factor=1.0, dq=None, dstar=None, diag=None):
        """pre + lower + upper on torch's current stream; returns (dq, dstar, diag)."""
        lib, N, F = self.lib, self.N, self.F
        self.check_inputs(rhsb, dt, fspr, dq=dq, dstar=dstar, diag=diag)
        diag = torch.empty(N, dtype=torch.float64, device=rhsb.device) if diag is None else diag
        dstar = torch.zeros_like(rhsb) if dstar is None else dstar
        dq = torch.zeros_like(rhsb) if dq is None else dq
        p = dict(nei=self.nei.data_ptr(), ic=self.icolor.data_ptr(), lc=self.lcolor.data_ptr(),
                 fnv=self.fnv.data_ptr(), vfn=self.vfn.data_ptr(), up=self.upts.data_ptr())
        with _Stream(lib):
            L.check(lib.parallel_pre_lusgs(N, F, float(factor), p["fnv"], dt.data_ptr(),
                                           diag.data_ptr(), fspr.data_ptr()), "pre")
            self._lower(p, rhsb, dstar, diag, fspr)
            for c in reversed(range(len(self.off) - 1)):
                L.check(lib.ns_parallel_upper_sweep(
                    self.off[c], self.off[c + 1], N, F, p["nei"], p["ic"], p["lc"], p["fnv"],
                    p["vfn"], p["up"], dq.data_ptr(), dstar.data_ptr(), diag.data_ptr(),
                    fspr.data_ptr()), "upper")
        return dq, dstar, diag

    def _lower(self, p, rhsb, dstar, diag, fspr):
        lib, N, F = self.lib, self.N, self.F
        for c in range(len(self.off) - 1):
            L.check(lib.ns_parallel_lower_sweep(
                self.off[c], self.off[c + 1], N, F, p["nei"], p["ic"], p["lc"], p["fnv"],
                p["vfn"], p["up"], rhsb.data_ptr(), dstar.data_ptr(), diag.data_ptr(),
                fspr.data_ptr()), "lower")

    def forward_inplace(self, rhsb, dt, fspr, factor=1.0):
        """pre + lower + upper with the in-place C sweeps on one buffer;
        returns dQ (dQ* is overwritten by the upper sweep)."""
        lib, N, F = self.lib, self.N, self.F
        self.check_inputs(rhsb, dt, fspr)
        diag = torch.empty(N, dtype=torch.float64, device=rhsb.device)
        dub = rhsb.clone()
        p = dict(nei=self.nei.data_ptr(), ic=self.icolor.data_ptr(), lc=self.lcolor.data_ptr(),
                 fnv=self.fnv.data_ptr(), vfn=self.vfn.data_ptr(), up=self.upts.data_ptr())
        with _Stream(lib):
            L.check(lib.parallel_pre_lusgs(N, F, float(factor), p["fnv"], dt.data_ptr(),
                                           diag.data_ptr(), fspr.data_ptr()), "pre")
            nc = len(self.off) - 1
            for c in range(nc):
                L.check(lib.parallel_ns_lower_sweep(
                    self.off[c], self.off[c + 1], N, F, p["nei"], p["ic"], p["lc"], p["fnv"],
                    p["vfn"], p["up"], dub.data_ptr(), diag.data_ptr(), fspr.data_ptr()), "lower")
            for c in reversed(range(nc)):
                L.check(lib.parallel_ns_upper_sweep(
                    self.off[c], self.off[c + 1], N, F, p["nei"], p["ic"], p["lc"], p["fnv"],
                    p["vfn"], p["up"], dub.data_ptr(), diag.data_ptr(), fspr.data_ptr()), "upper")
        return dub

    def recompute_dstar(self, rhsb, dt, fspr, factor=1.0):
        lib = self.lib
        self.check_inputs(rhsb, dt, fspr)
        diag = torch.empty(self.N, dtype=torch.float64, device=rhsb.device)
        dstar = torch.zeros_like(rhsb)
        p = dict(nei=self.nei.data_ptr(), ic=self.icolor.data_ptr(), lc=self.lcolor.data_ptr(),
                 fnv=self.fnv.data_ptr(), vfn=self.vfn.data_ptr(), up=self.upts.data_ptr())
        with _Stream(lib):
            L.check(lib.parallel_pre_lusgs(self.N, self.F, float(factor), p["fnv"],
                                           dt.data_ptr(), diag.data_ptr(), fspr.data_ptr()), "pre")
            self._lower(p, rhsb, dstar, diag, fspr)
        return dstar

    def adjoint(self, dt, fspr, dstar, dq, w, factor=1.0, overwrite_dstar=False):
        """(rhsbar, dtbar, fsprbar) of J = sum(w * dq); with overwrite_dstar
        the rhsb gradient is written into dstar's buffer and returned."""
        w = w.contiguous()
        self.check_inputs(None, dt, fspr, dq=dq, dstar=dstar, w=w)
        rb = dstar if overwrite_dstar else torch.zeros_like(dq)
        dtb = torch.empty_like(dt)
        fb = torch.empty_like(fspr)
        with _Stream(self.lib):
            L.check(self.lib.ucfd_lusgs_adjoint(self.handle, float(factor), dt.data_ptr(),
                                                fspr.data_ptr(), dstar.data_ptr(), dq.data_ptr(),
                                                w.data_ptr(), rb.data_ptr(), dtb.data_ptr(),
                                                fb.data_ptr()), "ucfd_lusgs_adjoint")
        return rb, dtb, fb

    def close(self):
        if self.handle:
            self.lib.ucfd_adjoint_destroy(self.handle)
            self.lib.ucfd_release_colored_layout(self.icolor.data_ptr(), self.lcolor.data_ptr())
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LUSGSPass(torch.autograd.Function):
    """dQ = coloured LU-SGS pass (pre + lower + upper) of rhsb [NFVARS, N],
    dt [N], fspr [F, N]; differentiable in all three."""

    @staticmethod
    def forward(ctx, rhsb, dt, fspr, system, recompute=False, factor=1.0):
        rhsb, dt, fspr = rhsb.contiguous(), dt.contiguous(), fspr.contiguous()
        system.check_inputs(rhsb, dt, fspr)
        ctx.system, ctx.recompute, ctx.factor = system, recompute, factor
        if recompute:
            dq = system.forward_inplace(rhsb, dt, fspr, factor)
            ctx.save_for_backward(rhsb, dt, fspr, dq)
        else:
            dq, dstar, _ = system.forward_arrays(rhsb, dt, fspr, factor)
            ctx.save_for_backward(rhsb, dt, fspr, dq, dstar)
        return dq

    @staticmethod
    def backward(ctx, w):
        S = ctx.system
        if ctx.recompute:
            rhsb, dt, fspr, dq = ctx.saved_tensors
            dstar = S.recompute_dstar(rhsb, dt, fspr, ctx.factor)
            rb, dtb, fb = S.adjoint(dt, fspr, dstar, dq, w, ctx.factor, overwrite_dstar=True)
        else:
            rhsb, dt, fspr, dq, dstar = ctx.saved_tensors
            rb, dtb, fb = S.adjoint(dt, fspr, dstar, dq, w, ctx.factor)
        return rb, dtb, fb, None, None, None


def lusgs_pass(system, rhsb, dt, fspr, recompute=False, factor=1.0):
    return LUSGSPass.apply(rhsb, dt, fspr, system, recompute, factor)

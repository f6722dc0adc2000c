"""Differentiation form on the GPU: the adjoint kernels (ucfd_lusgs_adjoint)
against the CPU adjoint oracle on the same inputs, and the
torch.autograd.Function (ucfd_sparse_amd.autograd.LUSGSPass) against central
differences of the GPU forward (h = 1e-6, FP64), with dQ* stored and
recomputed (identical gradients)."""
import math

import numpy as np
import pytest

from oracle import pyoracle as O
from ucfd_sparse_amd import layout as LY
from ucfd_sparse_amd import mesh as M

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SCALE = {"rhsb": 1.0, "dt": 1.0, "fspr": 1e3}   # see tests/test_adjoint_oracle.py


def say(msg):
    print(f"[adjoint-gpu] {msg}", flush=True)


@pytest.fixture(scope="module")
def setup():
    from ucfd_sparse_amd.autograd import ColoredSystem
    mesh = M.tet_mesh(5, seed=19, permute=True)
    flow = M.flow_fields(mesh, seed=19)
    lay = LY.colored_layout(mesh, flow)
    S = ColoredSystem(lay)
    yield lay, S
    S.close()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_adjoint_kernels_match_oracle(setup):
    lay, S = setup
    m, f = lay.mesh, lay.flow
    say(f"adjoint vs oracle, {m.neles} cells, {lay.ncolors} colours")
    rng = np.random.default_rng(5)
    w = rng.normal(size=(5, m.neles))
    rhs, dt, fspr = dev(f.rhsb[:5]), dev(f.dt), dev(f.fspr)
    dq, dstar, _ = S.forward_arrays(rhs, dt, fspr)
    torch.cuda.synchronize()
    od, ods, odq, _ = O.colored_pass(m.nei_ele, lay.icolor, lay.lcolor, m.fnorm_vol, m.vec_fnorm,
                                     f.fspr, f.dt, f.upts, f.rhsb)
    assert np.array_equal(dq.cpu().numpy(), odq)
    assert np.array_equal(dstar.cpu().numpy(), ods)
    rb, dtb, fb = S.adjoint(dt, fspr, dstar, dq, dev(w))
    torch.cuda.synchronize()
    orb, odtb, ofb = O.lusgs_adjoint(m.nei_ele, lay.off, m.fnorm_vol, m.vec_fnorm, f.fspr, f.dt,
                                     f.upts, ods, odq, w)
    for name, g, o in (("rhsbar", rb, orb), ("dtbar", dtb, odtb), ("fsprbar", fb, ofb)):
        g = g.cpu().numpy()
        diff = np.abs(g - o).max() / np.abs(o).max()
        say(f"  {name}: bitwise={np.array_equal(g, o)} max rel diff {diff:.1e}")
        assert diff <= 1e-13


@pytest.mark.parametrize("recompute", [False, True])
def test_autograd_matches_central_differences(setup, recompute):
    from ucfd_sparse_amd.autograd import lusgs_pass
    lay, S = setup
    m, f = lay.mesh, lay.flow
    say(f"autograd vs central differences, recompute={recompute}")
    rng = np.random.default_rng(7)
    w = dev(rng.normal(size=(5, m.neles)))
    rhs = dev(f.rhsb[:5]).requires_grad_()
    dt = dev(f.dt).requires_grad_()
    fspr = dev(f.fspr).requires_grad_()
    J = (w * lusgs_pass(S, rhs, dt, fspr, recompute=recompute)).sum()
    J.backward()
    grads = {"rhsb": rhs.grad, "dt": dt.grad, "fspr": fspr.grad}
    base = {"rhsb": rhs.detach(), "dt": dt.detach(), "fspr": fspr.detach()}

    def Jof(**kw):
        x = dict(base, **kw)
        dq, _, _ = S.forward_arrays(x["rhsb"], x["dt"], x["fspr"])
        return math.fsum((w * dq).double().cpu().numpy().ravel())
    h = 1e-6
    worst = {}
    for name in ("rhsb", "dt", "fspr"):
        for k in range(8):
            d = torch.from_numpy(rng.normal(size=tuple(base[name].shape)) * SCALE[name]).cuda()
            fd = (Jof(**{name: base[name] + h * d}) - Jof(**{name: base[name] - h * d})) / (2 * h)
            ad = float((grads[name] * d).sum())
            worst[name] = max(worst.get(name, 0.0), abs(fd - ad) / abs(ad))
    say(f"  worst relative FD disagreement {worst}")
    assert max(worst.values()) <= 1e-6
    if recompute:
        rhs2 = base["rhsb"].clone().requires_grad_()
        dt2 = base["dt"].clone().requires_grad_()
        fs2 = base["fspr"].clone().requires_grad_()
        (w * lusgs_pass(S, rhs2, dt2, fs2, recompute=False)).sum().backward()
        assert torch.equal(rhs2.grad, rhs.grad)
        assert torch.equal(dt2.grad, dt.grad)
        assert torch.equal(fs2.grad, fspr.grad)


def test_non_coloured_layout_is_rejected():
    import ctypes

    from ucfd_sparse_amd import lib as L
    mesh = M.tet_mesh(3, seed=3, permute=True)
    nei = dev(mesh.nei_ele)
    h = ctypes.c_void_p()
    off = np.array([0, mesh.neles], np.int32)
    lib = L.load()
    assert lib.ucfd_adjoint_create.argtypes is not None
    rc = lib.ucfd_adjoint_create(ctypes.byref(h), mesh.neles, mesh.nface, 1, off.ctypes.data,
                                 nei.data_ptr(), None, None, None)
    assert rc != 0

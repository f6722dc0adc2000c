"""ILU(0) family on the GPU, through the C ABI (UCFDPreconSetILU / SetBILU /
SetPBILU / SetCUDABILU, UCFDPreconPrepare, UCFDPreconApply, UCFDSolve),
against the CPU oracle (oracle/ilu_oracle.c) on the same seeded inputs and
against the reference's own outputs (tests/golden/ilu_tet3p.npz).

Bars: the in-place factorisations (level-scheduled; block rows shared by
the lanes of a wave, or one thread per row when a lower slot names its own
row) and the triangular solves are bit-exact (same operation order, no FMA
contraction); GMRES / BiCGSTAB preconditioned by them follow the oracle with
the bars of the other Krylov tests (same iteration counts and status,
histories within 1e-6, true residual within the tolerance).
"""
import ctypes
import os
import sys

import numpy as np
import pytest

from oracle import pyoracle as O
from ucfd_sparse_amd import layout as LY
from ucfd_sparse_amd import mesh as M

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden  # noqa: E402

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

FIX = os.path.join(HERE, "golden", "ilu_tet3p.npz")


def say(msg):
    print(f"[ilu-gpu] {msg}", flush=True)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


class Dev:
    """Device copy of a BSR matrix (the preconditioner factors its values)."""

    def __init__(self, A):
        self.bn, self.blk = A.bn, A.blk
        self.rowptr, self.colidx, self.diagslots = dev(A.rowptr), dev(A.colidx), dev(A.diagslots)
        self.values = dev(A.values)


class Host:
    """Host arrays of a BSR matrix with a private copy of the values."""

    def __init__(self, A):
        self.bn, self.blk = A.bn, A.blk
        self.rowptr, self.colidx, self.diagslots = A.rowptr, A.colidx, A.diagslots
        self.values = np.array(A.values, np.float64)


@pytest.fixture(scope="module")
def K():
    from ucfd_sparse_amd import krylov
    return krylov


@pytest.fixture(scope="module")
def cases():
    return make_golden.ilu_cases()


def pkind(kind):
    return "ilu" if kind == "ilu" else "bilu"


def test_reference_cases_bitwise_host_arrays(K, cases):
    fix = np.load(FIX)
    for name, kind, A, ic, b in cases:
        say(f"{name}: bn={A.bn} blk={A.blk} ({kind}, host arrays)")
        H = Host(A)
        pc = K.Precon.from_bsr(H, icolors=ic if kind == "pbilu" else None, kind=pkind(kind))
        f = O.ilu_prepare(A, kind)
        assert np.array_equal(H.values, f.reshape(H.values.shape)), name
        assert np.array_equal(H.values.reshape(-1), fix[f"{name}_factored"].reshape(-1)), name
        x = np.array(b, np.float64)
        pc.apply(x)
        assert np.array_equal(x, O.ilu_apply(A, f, b, kind, ic)), name
        if kind != "ilu":
            assert np.array_equal(x, fix[f"{name}_apply"]), name


def random_case(bn, blk, per_row, seed, ncolors=0):
    r = make_golden.random_pattern_bsr(bn, blk, per_row, seed, ncolors=ncolors)
    A, off = r if ncolors else (r, None)
    return A, off, M.normal_fill(A.bn * A.blk, seed + 1)


@pytest.mark.parametrize("blk", [1, 2, 3, 4, 5, 6, 7])
def test_block_ilu_random_patterns_device_bitwise(K, blk):
    """Random symmetric patterns with triangles (fill inside the pattern),
    sequential form: many dependency levels, ordered (gathered) levels."""
    A, _, b = random_case(600, blk, 7, 500 + blk)
    say(f"block ILU(0) random pattern bn={A.bn} blk={blk} (device arrays)")
    D = Dev(A)
    pc = K.Precon.from_bsr(D, kind="bilu")
    f = O.ilu_prepare(A, "bilu")
    assert np.array_equal(D.values.cpu().numpy(), f.reshape(A.values.shape))
    x = dev(b)
    pc.apply(x)
    assert np.array_equal(x.cpu().numpy(), O.ilu_apply(A, f, b, "bilu"))


def test_scalar_ilu_device_bitwise(K):
    A, _, b = random_case(3000, 1, 9, 611)
    say(f"scalar ILU(0) random pattern n={A.bn} (device arrays)")
    D = Dev(A)
    pc = K.Precon(D.rowptr, D.colidx, D.diagslots, D.values, A.bn, 1, kind="ilu")
    f = O.ilu_prepare(A, "ilu")
    assert np.array_equal(D.values.cpu().numpy(), f.reshape(A.values.shape))
    x = dev(b)
    pc.apply(x)
    assert np.array_equal(x.cpu().numpy(), O.ilu_apply(A, f, b, "ilu"))


@pytest.mark.parametrize("blk", [1, 5])
def test_colored_block_ilu_tet_mesh_device_bitwise(K, blk):
    """Colour-ordered tet mesh: one launch per colour range for the
    factorisation and for each sweep."""
    mesh = M.tet_mesh(10, seed=620 + blk, permute=True)
    flow = M.flow_fields(mesh, seed=620 + blk)
    lay = LY.colored_layout(mesh, flow, coloring="greedy-bfs")
    A = M.block_system(lay.mesh, lay.flow, blk=blk, seed=621)
    b = M.normal_fill(A.bn * blk, 622)
    say(f"coloured block ILU(0) tet mesh bn={A.bn} blk={blk} colours={len(lay.off) - 1}")
    D = Dev(A)
    pc = K.Precon.from_bsr(D, icolors=lay.off, kind="bilu")
    f = O.ilu_prepare(A, "pbilu")
    assert np.array_equal(D.values.cpu().numpy(), f.reshape(A.values.shape))
    x = dev(b)
    pc.apply(x)
    ref = O.ilu_apply(A, f, b, "pbilu", lay.off)
    assert np.array_equal(x.cpu().numpy(), ref)
    # the same numbering without colours: the sequential form gives the same result
    D2 = Dev(A)
    pc2 = K.Precon.from_bsr(D2, kind="bilu")
    assert torch.equal(D2.values, D.values)
    x2 = dev(b)
    pc2.apply(x2)
    assert np.array_equal(x2.cpu().numpy(), ref)


def test_colored_block_ilu_with_coupled_ranges_bitwise(K):
    """Colour ranges that couple to themselves: levels are built inside the
    reference's visit order (ranges ascending / descending, rows ascending)."""
    A, _, b = random_case(900, 3, 6, 630)
    off = np.array([0, 200, 650, 900], np.int32)
    say(f"coloured block ILU(0), coupled ranges, bn={A.bn}")
    D = Dev(A)
    pc = K.Precon.from_bsr(D, icolors=off, kind="bilu")
    f = O.ilu_prepare(A, "pbilu")
    assert np.array_equal(D.values.cpu().numpy(), f.reshape(A.values.shape))
    x = dev(b)
    pc.apply(x)
    assert np.array_equal(x.cpu().numpy(), O.ilu_apply(A, f, b, "pbilu", off))


@pytest.mark.parametrize("blk,ncolors", [(1, 0), (3, 0), (5, 0), (4, 3)])
def test_self_slots_bitwise(K, blk, ncolors):
    """Lower slots whose column is the row itself: the block factorisation
    leaves the wave-per-rows kernel for the thread-per-row one (a row's
    blocks are then read while the row is written), the solves read the
    row's own partial result; all as the sequential reference does."""
    A0, off, b = random_case(500, blk, 6, 680 + blk, ncolors=ncolors)
    A = make_golden.with_self_slots(A0, 4, 681)
    kind = "ilu" if blk == 1 else ("pbilu" if ncolors else "bilu")
    say(f"{kind} with self slots bn={A.bn} blk={blk}")
    D = Dev(A)
    if blk == 1:
        pc = K.Precon(D.rowptr, D.colidx, D.diagslots, D.values, A.bn, 1, kind="ilu")
    else:
        pc = K.Precon.from_bsr(D, icolors=off, kind="bilu")
    f = O.ilu_prepare(A, kind)
    assert np.array_equal(D.values.cpu().numpy(), f.reshape(A.values.shape))
    x = dev(b)
    pc.apply(x)
    assert np.array_equal(x.cpu().numpy(), O.ilu_apply(A, f, b, kind, off))


def test_second_prepare_factors_again(K):
    """The factorisation works in place on the caller's values (as the
    reference's does): a second Prepare factors the factored values."""
    A, _, b = random_case(300, 2, 5, 640)
    say("second Prepare on the factored values")
    D = Dev(A)
    pc = K.Precon.from_bsr(D, kind="bilu")
    pc.prepare()
    f1 = O.ilu_prepare(A, "bilu")
    A2 = M.BSR(A.bn, A.blk, A.rowptr, A.colidx, A.diagslots, f1.reshape(A.values.shape))
    f2 = O.ilu_prepare(A2, "bilu")
    assert np.array_equal(D.values.cpu().numpy(), f2.reshape(A.values.shape))


def test_cuda_named_setter_is_the_coloured_form(K):
    from ucfd_sparse_amd import lib as L
    lib = L.load()
    A, off, b = random_case(400, 4, 6, 650, ncolors=3)
    say("UCFDPreconSetCUDABILU")
    D = Dev(A)
    h = ctypes.c_void_p()
    L.check(lib.UCFDPreconCreatefromArrays(ctypes.byref(h), L.ptr(D.rowptr), L.ptr(D.colidx),
                                           L.ptr(D.diagslots), L.ptr(D.values)),
            "UCFDPreconCreatefromArrays")
    ic = np.ascontiguousarray(off, np.int32)
    try:
        L.check(lib.UCFDPreconSetCUDABILU(ctypes.byref(h), A.bn, A.blk, len(ic) - 1,
                                          ic.ctypes.data), "UCFDPreconSetCUDABILU")
        L.check(lib.UCFDPreconPrepare(h), "UCFDPreconPrepare")
        x = dev(b)
        L.check(lib.UCFDPreconApply(h, L.ptr(x)), "UCFDPreconApply")
    finally:
        lib.UCFDPreconDestroy(ctypes.byref(h))
    f = O.ilu_prepare(A, "pbilu")
    assert np.array_equal(x.cpu().numpy(), O.ilu_apply(A, f, b, "pbilu", off))


def _close_hist(h_gpu, h_orc, n):
    h_gpu, h_orc = np.asarray(h_gpu[:n]), np.asarray(h_orc[:n])
    if n == 0:
        return 0.0
    scale = np.maximum(np.abs(h_orc), 1e-6 * abs(h_orc[0]))
    return float((np.abs(h_gpu - h_orc) / scale).max())


@pytest.mark.parametrize("solver", ["gmres", "bicgstab"])
@pytest.mark.parametrize("colored", [False, True])
def test_preconditioned_solves_match_oracle(K, solver, colored):
    mesh = M.tet_mesh(6, seed=660, permute=True)
    flow = M.flow_fields(mesh, seed=660)
    off = None
    if colored:
        lay = LY.colored_layout(mesh, flow, coloring="greedy-bfs")
        mesh, flow, off = lay.mesh, lay.flow, lay.off
    A = M.block_system(mesh, flow, blk=5, seed=661)
    b = M.normal_fill(A.bn * A.blk, 662)
    kind = "pbilu" if colored else "bilu"
    tol = 1e-11
    say(f"{solver} + {kind}, bn={A.bn}")
    if solver == "gmres":
        m, maxiter = 5, 40
        xo, ito, sto, reso, histo = O.gmres_ilu(A, b, kind, m, maxiter, tol, icolors=off)
    else:
        maxiter = 60
        xo, ito, sto, reso, histo = O.bicgstab_ilu(A, b, kind, maxiter, tol, icolors=off)
    D = Dev(A)
    mat = K.SpMat.from_bsr(Dev(A))
    pc = K.Precon.from_bsr(D, icolors=off, kind="bilu")
    s = K.Solver(solver, A.bn * A.blk, maxiter, tol, m=5)
    x = dev(np.zeros(A.bn * A.blk))
    st, it, res = s.solve(pc, mat, x, dev(b))
    say(f"  gpu it={it} st={st} res={res:.3e}; oracle it={ito} st={sto} res={reso:.3e}")
    assert (it, st) == (ito, sto)
    n = it + 1 if it < maxiter else it
    assert _close_hist(s.history(), histo, n) <= 1e-6
    xg = x.cpu().numpy()
    assert np.linalg.norm(b - A.to_scipy() @ xg) <= 10 * tol
    assert np.allclose(xg, xo, rtol=1e-6, atol=1e-9 * np.abs(xo).max())


def test_errors_are_reported(K):
    from ucfd_sparse_amd import lib as L
    A, _, b = random_case(50, 1, 4, 670)
    say("ILU errors: zero pivot, mixed precision")
    vals = np.array(A.values, np.float64)
    vals[A.diagslots[0]] = 0.0          # row 0 has no lower entries: its pivot stays 0
    with pytest.raises(L.UCFDError):
        K.Precon(A.rowptr, A.colidx, A.diagslots, vals, A.bn, 1, kind="ilu")
    with pytest.raises(L.UCFDError):
        K.Precon(A.rowptr, A.colidx, A.diagslots, np.array(A.values), A.bn, 1, kind="bilu",
                 fp32=True)

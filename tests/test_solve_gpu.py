"""GPU defect-correction solver, FP64 and mixed FP32/FP64 forms, against the
oracle (tests/test_solve_oracle.py pins the oracle's operator).

Tolerances: the GPU evaluates the flux difference in the exact difference
form while the oracle subtracts two fluxes (lusgs.c:143-155), whose
cancellation costs ~1e-11 relative for |dq| ~ 1e-5 |Q|; residuals are
therefore compared to 1e-10 of ||b||_inf, and converged solutions to 1e-8
relative (both are converged to 1e-10 in the residual)."""
import numpy as np
import pytest

import pyoracle as O
from ucfd_sparse_amd import lib as L
from ucfd_sparse_amd import mesh as M
from ucfd_sparse_amd.layout import colored_layout
from ucfd_sparse_amd.solver import LUSGSSolver

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if L.device_count() == 0:
        pytest.fail("no GPU visible")
    L.call("ucfd_set_config", 5, 5, 3)
    L.call("ucfd_set_stream", None)


def make(m, seed, hexmesh=False, coloring="greedy"):
    mesh = M.hex_mesh(m, m, m, seed=seed) if hexmesh else M.tet_mesh(m, seed=seed)
    flow = M.flow_fields(mesh, seed=seed)
    return colored_layout(mesh, flow, coloring=coloring)


# greedy: 4-5 colours; greedy-bfs: 2 colours on these meshes (the first
# colour's lower step and the last colour's upper step then sit in the
# other colour's launch, as in the 96^3 x 6 measurement)
SOLVE_CASES = [(6, 3, False), (8, 5, True), (12, 4, False), (6, 3, False, "greedy-bfs"),
               (8, 5, True, "greedy-bfs")]


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


@pytest.mark.parametrize("case", [(6, 3, False), (8, 5, True)])
def test_residual_matches_oracle(case):
    lay = make(*case)
    m, f = lay.mesh, lay.flow
    s = LUSGSSolver.from_layout(lay)
    dq = np.random.default_rng(2).normal(scale=1e-5, size=(5, m.neles))
    r = np.zeros_like(dq)
    nrm = s.residual(np.ascontiguousarray(f.rhsb[:5]), dq, r)
    ro, no = O.residual(m.nei_ele, m.fnorm_vol, m.vec_fnorm, f.fspr, f.dt, f.upts, f.rhsb, dq)
    e = np.abs(r - ro).max() / np.abs(f.rhsb).max()
    say(f"residual vs oracle: {e:.2e}")
    assert e <= 1e-10
    np.testing.assert_allclose(nrm, no, rtol=1e-9)
    # r(0) = b exactly
    r0 = np.zeros_like(dq)
    s.residual(np.ascontiguousarray(f.rhsb[:5]), np.zeros_like(dq), r0)
    assert np.array_equal(r0, f.rhsb[:5])
    s.close()


def test_gpu_residual_vanishes_on_first_colour_after_legacy_pass():
    """Ties the GPU operator to the GPU legacy sweeps (bit-exact with the
    reference): after one pass the colour-1 residual is rounding only."""
    lay = make(7, 9)
    m, f = lay.mesh, lay.flow
    N = m.neles
    diag = np.zeros(N)
    dub = np.ascontiguousarray(f.rhsb[:5]).copy()
    L.call("parallel_pre_lusgs", N, 4, 1.0, m.fnorm_vol, f.dt, diag, f.fspr)
    off = [int(x) for x in lay.off]
    for c in range(lay.ncolors):
        L.call("parallel_ns_lower_sweep", off[c], off[c + 1], N, 4, m.nei_ele, lay.icolor,
               lay.lcolor, m.fnorm_vol, m.vec_fnorm, f.upts, dub, diag, f.fspr)
    for c in reversed(range(lay.ncolors)):
        L.call("parallel_ns_upper_sweep", off[c], off[c + 1], N, 4, m.nei_ele, lay.icolor,
               lay.lcolor, m.fnorm_vol, m.vec_fnorm, f.upts, dub, diag, f.fspr)
    s = LUSGSSolver.from_layout(lay)
    r = np.zeros_like(dub)
    s.residual(np.ascontiguousarray(f.rhsb[:5]), dub, r)
    c1 = lay.lcolor == 1
    scale = np.abs(f.rhsb).max()
    say(f"colour-1 residual {np.abs(r[:, c1]).max() / scale:.2e}")
    assert np.abs(r[:, c1]).max() <= 1e-11 * scale
    s.close()


@pytest.mark.parametrize("case", SOLVE_CASES)
def test_fp64_solve_matches_oracle_defect_correction(case):
    lay = make(*case)
    if len(case) > 3:
        assert lay.ncolors == 2
    m, f = lay.mesh, lay.flow
    s = LUSGSSolver.from_layout(lay)
    b = np.ascontiguousarray(f.rhsb[:5])
    dq = np.zeros_like(b)
    st = s.solve(b, dq, "fp64", 1e-10, 100)
    dqo, npass, hist = O.dc_solve(m.nei_ele, lay.icolor, lay.lcolor, m.fnorm_vol, m.vec_fnorm,
                                  f.fspr, f.dt, f.upts, f.rhsb, 1e-10, 100)
    e = np.abs(dq - dqo).max() / np.abs(dqo).max()
    say(f"fp64 passes gpu {st.passes} oracle {npass}; rel res {st.rel_residual:.2e}; "
        f"dq vs oracle {e:.2e}")
    assert st.converged and st.rel_residual <= 1e-10
    assert abs(st.passes - npass) <= 1
    assert e <= 1e-8
    ro, no = O.residual(m.nei_ele, m.fnorm_vol, m.vec_fnorm, f.fspr, f.dt, f.upts, f.rhsb, dq)
    assert no <= 2e-10 * np.linalg.norm(b)
    s.close()


@pytest.mark.parametrize("case", SOLVE_CASES)
def test_mixed_solve_reaches_fp64_accuracy(case):
    lay = make(*case)
    if len(case) > 3:
        assert lay.ncolors == 2
    m, f = lay.mesh, lay.flow
    s = LUSGSSolver.from_layout(lay)
    b = dev(f.rhsb[:5])
    dq64 = dev(np.zeros((5, m.neles)))
    dqm = dev(np.zeros((5, m.neles)))
    s64 = s.solve(b, dq64, "fp64", 1e-10, 100)
    sm = s.solve(b, dqm, "mixed", 1e-10, 200)
    a64, am = dq64.cpu().numpy(), dqm.cpu().numpy()
    e = np.abs(am - a64).max() / np.abs(a64).max()
    say(f"mixed: outer {sm.outer} passes {sm.passes} rel res {sm.rel_residual:.2e}; "
        f"fp64 passes {s64.passes}; error vs fp64 {e:.2e}; hist {sm.hist()}")
    assert sm.converged and sm.rel_residual <= 1e-10
    assert e <= 1e-8
    ro, no = O.residual(m.nei_ele, m.fnorm_vol, m.vec_fnorm, f.fspr, f.dt, f.upts, f.rhsb, am)
    assert no <= 2e-10 * np.linalg.norm(f.rhsb[:5])
    s.close()


def test_solver_rejects_invalid_colouring():
    lay = make(4, 1)
    m, f = lay.mesh, lay.flow
    bad = lay.off.copy()
    bad[1] = bad[2]           # merge colours 1 and 2: adjacent cells share a "colour"
    with pytest.raises(L.UCFDError):
        LUSGSSolver(m.neles, 4, bad, m.nei_ele, m.fnorm_vol, m.vec_fnorm, f.fspr, f.dt,
                    np.ascontiguousarray(f.upts[:5]))

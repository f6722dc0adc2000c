"""Pins of the oracle's LU-SGS system operator (oracle/residual_oracle.c) and
of its FP64 defect correction.  The reference has no residual routine, so
the operator is pinned by what the mathematics fixes:
  * r(0) = b exactly;
  * on an all-self (all-boundary) mesh r = b - D dq, D the pre diagonal;
  * after one coloured LU-SGS pass (the oracle's sweeps, themselves pinned
    bit for bit to the compiled reference) the residual vanishes on colour 1:
    there the two sweeps together are exactly the residual's own equation, so
    a wrong sign, factor or face set in the operator's neighbour term fails;
  * the defect correction drives the residual below rtol.
"""
import numpy as np

import pyoracle as O
from ucfd_sparse_amd import mesh as M
from ucfd_sparse_amd.layout import colored_layout


def setup(m=5, seed=3):
    mesh = M.tet_mesh(m, seed=seed)
    flow = M.flow_fields(mesh, seed=seed)
    lay = colored_layout(mesh, flow)
    return lay


def test_residual_at_zero_is_rhs():
    lay = setup()
    m, f = lay.mesh, lay.flow
    r, nrm = O.residual(m.nei_ele, m.fnorm_vol, m.vec_fnorm, f.fspr, f.dt, f.upts, f.rhsb,
                        np.zeros_like(f.rhsb))
    assert np.array_equal(r, f.rhsb)
    np.testing.assert_allclose(nrm, np.linalg.norm(f.rhsb), rtol=1e-14)


def test_residual_all_self_mesh_is_diagonal():
    lay = setup()
    m, f = lay.mesh, lay.flow
    N = m.neles
    self_nei = np.tile(np.arange(N, dtype=np.int32), (4, 1))
    dq = np.random.default_rng(1).normal(scale=1e-5, size=f.rhsb.shape)
    r, _ = O.residual(self_nei, m.fnorm_vol, m.vec_fnorm, f.fspr, f.dt, f.upts, f.rhsb, dq)
    D = 1.0 / f.dt + 0.5 * (f.fspr * m.fnorm_vol).sum(0)
    np.testing.assert_allclose(r, f.rhsb - D * dq, rtol=1e-13, atol=1e-18)


def test_residual_vanishes_on_first_colour_after_one_pass():
    lay = setup(6, 7)
    m, f = lay.mesh, lay.flow
    _, _, dq, _ = O.colored_pass(m.nei_ele, lay.icolor, lay.lcolor, m.fnorm_vol, m.vec_fnorm,
                                 f.fspr, f.dt, f.upts, f.rhsb)
    r, _ = O.residual(m.nei_ele, m.fnorm_vol, m.vec_fnorm, f.fspr, f.dt, f.upts, f.rhsb, dq)
    c1 = lay.lcolor == 1
    scale = np.abs(f.rhsb).max()
    assert np.abs(r[:, c1]).max() <= 1e-12 * scale
    # and it does not vanish elsewhere (the operator is not the sweep)
    assert np.abs(r[:, ~c1]).max() > 1e-8 * scale


def test_defect_correction_converges():
    lay = setup(6, 11)
    m, f = lay.mesh, lay.flow
    dq, n, hist = O.dc_solve(m.nei_ele, lay.icolor, lay.lcolor, m.fnorm_vol, m.vec_fnorm,
                             f.fspr, f.dt, f.upts, f.rhsb, 1e-10, 100)
    assert hist[-1] <= 1e-10 and n < 100
    assert np.all(np.diff(hist) < 0)
    r, nrm = O.residual(m.nei_ele, m.fnorm_vol, m.vec_fnorm, f.fspr, f.dt, f.upts, f.rhsb, dq)
    assert nrm <= 1e-10 * np.linalg.norm(f.rhsb[:5])

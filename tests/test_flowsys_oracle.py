"""The FlowSys block LU-SGS oracle (oracle/blusgs_oracle.c), pinned on the CPU:

* bit for bit against the reference's own outputs at every stage of an outer
  step (tests/golden/flowsys_tet3p.npz, written by tests/golden/
  make_golden.py --flowsys from the reference compiled out of src/lusgs/*.c
  and driven through its FlowSys API), for the Spalart-Allmaras and the
  k-omega SST turbulence packs, and live against that build where it exists;
* against closed forms: one lower + upper pass from du = 0 solves
  (D + L) D^-1 (D + U) du = rhs with D, L, U the block diagonal and the
  strictly lower / upper neighbour blocks the positions define; repeated
  passes converge to the solution of (D + L + U) du = rhs; the assembled
  diagonal block is (1/dt + a0) I + sum of the signed face Jacobians * fv.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from oracle import pyoracle as O  # noqa: E402
from ucfd_sparse_amd import mesh as M  # noqa: E402

import make_golden  # noqa: E402
import refflowsys  # noqa: E402

FIX = os.path.join(HERE, "golden", "flowsys_tet3p.npz")


@pytest.fixture(scope="module")
def case():
    return make_golden.flowsys_case()


@pytest.fixture(scope="module")
def fix():
    return np.load(FIX)


def test_fixture_inputs_regenerate(case, fix):
    mesh, flow, fs, tj, el = case
    d = make_golden.input_digest(fs.rowptr, fs.colidx, fs.slots, fs.sides, fs.face_area,
                                 fs.rcp_vol, fs.jmat, tj, *el.values())
    assert d == str(fix["digest"])


def test_every_stage_bitwise_vs_reference(case, fix):
    mesh, flow, fs, tj, el = case
    out = O.flowsys_run(fs, tj, el, 6, 5)
    for k, v in out.items():
        assert np.array_equal(v, fix[k]), k


def test_kwsst_pack_bitwise_vs_reference(case, fix):
    mesh, flow, fs, tj, el = case
    tj2 = M.flowsys_system(mesh, flow, nv=2, seed=303).jmat
    el2 = M.flowsys_elements(mesh, flow, nvars=7, seed=303)
    out = O.flowsys_run(fs, tj2, el2, 7, 5, kind="kwsst", iterations=1)
    assert np.array_equal(out["tdiag_pack"], fix["kwsst_tdiag_pack"])
    assert np.array_equal(out["du_rans0"], fix["kwsst_du_rans0"])


def blocks(fs, diag_pack, nv=5):
    """Dense D (unfactored), L and U of the NS block system."""
    N = fs.nlocal
    n = N * nv
    D = np.zeros((n, n))
    L = np.zeros((n, n))
    U = np.zeros((n, n))
    dp = diag_pack.reshape(N, nv, nv)
    for r in range(N):
        D[r * nv:(r + 1) * nv, r * nv:(r + 1) * nv] = dp[r]
        for p in range(fs.rowptr[r], fs.rowptr[r + 1]):
            s, nb = fs.slots[p], fs.colidx[p]
            fv = fs.face_area[s] * fs.rcp_vol[r]
            a, b = fs.jmat[0, :, :, s], fs.jmat[1, :, :, s]
            v, j = (a, b) if fs.sides[p] == 1 else (-b, -a)
            D[r * nv:(r + 1) * nv, r * nv:(r + 1) * nv] += v * fv
            T = L if nb < r else U
            T[r * nv:(r + 1) * nv, nb * nv:(nb + 1) * nv] += j * fv
    return D, L, U


def test_one_pass_is_the_sgs_factorisation(case):
    mesh, flow, fs, tj, el = case
    S = O.FlowSysOracle(fs, fs.jmat, None, el, 5, 5)
    S.el = dict(el, rhs=el["rhs"][:5], uptsb=el["uptsb"][:5])
    S.upts = np.array(el["uptsb"][:5])
    S.pack(0.3)
    D, L, U = blocks(fs, S.diag.copy())
    # the assembled block equals D, and its LU solves with D
    S.prepare()
    S.reset()
    S.sweep(False)
    S.sweep(True)
    rhs = S.rhs.copy()
    Mm = (D + L) @ np.linalg.solve(D, D + U)
    assert np.linalg.norm(Mm @ S.du - rhs) <= 1e-12 * np.linalg.norm(rhs)


def test_gauss_seidel_converges_to_the_block_system(case):
    mesh, flow, fs, tj, el = case
    S = O.FlowSysOracle(fs, fs.jmat, None, el, 5, 5)
    S.el = dict(el, rhs=el["rhs"][:5], uptsb=el["uptsb"][:5])
    S.upts = np.array(el["uptsb"][:5])
    S.pack(0.3)
    D, L, U = blocks(fs, S.diag.copy())
    S.prepare()
    S.reset()
    for _ in range(60):
        S.sweep(False)
        S.sweep(True)
    A = D + L + U
    r = A @ S.du - S.rhs
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(S.rhs)


@pytest.mark.skipif(not refflowsys.available(), reason="reference build absent")
def test_live_reference_on_a_colour_ordered_hex_mesh():
    from ucfd_sparse_amd.layout import colored_layout
    mesh = M.hex_mesh(5, 4, 3, seed=61, permute=True)
    flow = M.flow_fields(mesh, seed=61)
    lay = colored_layout(mesh, flow)
    fs = M.flowsys_system(lay.mesh, lay.flow, nv=5, seed=61)
    tj = M.flowsys_system(lay.mesh, lay.flow, nv=1, seed=61).jmat
    el = M.flowsys_elements(lay.mesh, lay.flow, nvars=6, seed=61)
    ref = refflowsys.run(fs, tj, el, 6, 5, a0=0.0, turb_factor=2.0, iterations=3)
    ours = O.flowsys_run(fs, tj, el, 6, 5, a0=0.0, turb_factor=2.0, iterations=3)
    for k in ref:
        assert np.array_equal(ours[k], ref[k]), k


# ---------------------------------------------------------------- scalar LU-SGS
# oracle/flowsys_lusgs_oracle.c, pinned bit for bit against the reference's
# FlowSys scalar LU-SGS (tests/golden/flowsys_lusgs_tet3p.npz, written by
# make_golden.py --flowsys-lusgs), live against the reference build where it
# exists, and against the legacy oracle through the renumbering bridge.
FIXL = os.path.join(HERE, "golden", "flowsys_lusgs_tet3p.npz")


@pytest.fixture(scope="module")
def lcase():
    return make_golden.flowsys_lusgs_case()


@pytest.fixture(scope="module")
def lfix():
    return np.load(FIXL)


def test_lusgs_fixture_inputs_regenerate(lcase, lfix):
    mesh, flow, fs, fspr, tfspr, el = lcase
    d = make_golden.input_digest(fs.rowptr, fs.colidx, fs.slots, fs.sides, fs.face_area,
                                 fs.face_normal, fs.rcp_vol, fspr, tfspr, *el.values())
    assert d == str(lfix["digest"])


def test_lusgs_every_stage_bitwise_vs_reference(lcase, lfix):
    mesh, flow, fs, fspr, tfspr, el = lcase
    out = O.flowsys_lusgs_run(fs, fspr, tfspr, el, 6, 5)
    assert sorted(out) == sorted(k for k in lfix.files if k != "digest")
    for k, v in out.items():
        assert np.array_equal(v, lfix[k]), k


def test_lusgs_case_meets_the_pressure_floor(lcase):
    """Some neighbour states of the NS sweeps have p < PMIN (the clamped
    branch of flux.c:50-56 is part of what the fixture pins)."""
    mesh, flow, fs, fspr, tfspr, el = lcase
    up = el["uptsb"]
    p = 0.4 * (up[4] - 0.5 * (up[1] ** 2 + up[2] ** 2 + up[3] ** 2) / up[0])
    assert (p < 1e-13).sum() >= 10


def test_lusgs_bridge_equals_the_legacy_oracle():
    """With one slot per (cell, face), sides +1, face_area = fnorm_vol,
    rcp_vol = 1, kappa = 1, a0 = dsrc = 0 and no turbulence rows, the FlowSys
    scalar sweeps are the legacy serial sweeps (two independent restatements
    of lusgs.c)."""
    mesh = M.tet_mesh(4, seed=41, permute=True)
    flow = M.flow_fields(mesh, seed=41)
    N, F = mesh.neles, mesh.nface
    nei = np.ascontiguousarray(mesh.nei_ele)
    rowptr = np.arange(N + 1, dtype=np.int32) * F
    fs = M.FlowSysMesh(N, F * N, rowptr, np.ascontiguousarray(nei.T.reshape(-1), np.int32),
                       np.arange(F * N, dtype=np.int32), np.ones(F * N, np.int8),
                       np.ascontiguousarray(mesh.fnorm_vol.T.reshape(-1)),
                       np.ascontiguousarray(np.stack([mesh.vec_fnorm[:, d, :].T.reshape(-1)
                                                      for d in range(3)])),
                       np.ones(N), None)
    lam = np.ascontiguousarray(flow.fspr.T.reshape(-1))
    el = {"cell_ids": np.arange(N, dtype=np.int32), "uptsb": flow.upts.copy(),
          "rhs": flow.rhsb.copy(), "dt": flow.dt.copy(), "dsrc": np.zeros_like(flow.upts),
          "vol": np.ones(N)}
    S = O.FlowSysLUSGSOracle(fs, lam, None, el, 5, 5)
    S.pack(1.0, 0.0)
    S.prepare(1.0)
    S.sweep(False, 1.0)
    dstar = S.du.reshape(5, N).copy()
    S.sweep(True, 1.0)
    dq = S.du.reshape(5, N).copy()
    diag, ds, dqo, _ = O.serial_pass(nei, mesh.fnorm_vol, mesh.vec_fnorm, flow.fspr, flow.dt,
                                     flow.upts, flow.rhsb)
    assert np.array_equal(S.diag.reshape(5, N)[0], diag)
    assert np.array_equal(dstar, ds)
    assert np.array_equal(dq, dqo)


@pytest.mark.skipif(not refflowsys.available(), reason="reference build absent")
def test_lusgs_live_reference_colour_ordered():
    mesh, flow, fs, fspr, tfspr, el = make_golden.flowsys_lusgs_case(seed=305, m=4, colored=True)
    ref = refflowsys.run_lusgs(fs, fspr, tfspr, el, 6, 5, a0=0.1, turb_factor=2.0, kappa=0.8,
                               iterations=3)
    out = O.flowsys_lusgs_run(fs, fspr, tfspr, el, 6, 5, a0=0.1, turb_factor=2.0, kappa=0.8,
                              iterations=3)
    for k, v in ref.items():
        assert np.array_equal(out[k], v), k

"""The incomplete-LU oracle (oracle/ilu_oracle.c), pinned on the CPU:

* factorisations (ILU(0), block ILU(0), coloured block ILU(0)) bit for bit
  against the reference's own outputs (tests/golden/ilu_tet3p.npz, written by
  tests/golden/make_golden.py --ilu from the reference compiled out of
  src/krylov and driven through UCFDPreconSetILU / SetBILU / SetPBILU +
  UCFDPreconPrepare), on tet-mesh matrices and on random symmetric patterns
  with triangles (fill updates inside the pattern);
* block applies bit for bit against the reference's apply, and GMRES /
  BiCGSTAB preconditioned by them (histories and solutions);
* the scalar apply: the reference's backward sweep starts at the diagonal
  entry (src/krylov/precon/ilu.c:81); its output is reproduced exactly by
  this oracle's sweep plus that one extra term, and this oracle's apply is
  the exact solve with the factors;
* closed forms: ILU(0) reproduces A on its pattern, (L U)_ij = A_ij; the
  applies solve L U x = b.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from oracle import pyoracle as O  # noqa: E402

import make_golden  # noqa: E402
import refkrylov  # noqa: E402

FIX = os.path.join(HERE, "golden", "ilu_tet3p.npz")


@pytest.fixture(scope="module")
def cases():
    return make_golden.ilu_cases()


@pytest.fixture(scope="module")
def fix():
    return np.load(FIX)


def dense_factors(A, f, kind):
    """Dense unit-lower L and upper U of the factored values (diagonal blocks
    rebuilt from their in-place LU for the block forms; a 1x1 block form keeps
    the reciprocal of its pivot)."""
    bn, blk = A.bn, A.blk
    n = bn * blk
    Lm, Um = np.eye(n), np.zeros((n, n))
    fv = f.reshape(-1, blk, blk)
    for i in range(bn):
        for p in range(A.rowptr[i], A.rowptr[i + 1]):
            c = A.colidx[p]
            blkv = fv[p]
            if p == A.diagslots[i] and blk > 1:
                lo = np.tril(blkv, -1) + np.eye(blk)
                blkv = lo @ np.triu(blkv)
            elif p == A.diagslots[i] and blk == 1 and kind != "ilu":
                blkv = 1.0 / blkv
            T = Lm if c < i else Um
            T[i * blk:(i + 1) * blk, c * blk:(c + 1) * blk] = blkv
    return Lm, Um


def test_inputs_regenerate(cases, fix):
    for name, kind, A, ic, b in cases:
        d = make_golden.input_digest(A.rowptr, A.colidx, A.diagslots, A.values, b,
                                     *([] if ic is None else [ic]))
        assert d == str(fix[f"{name}_digest"]), name


def test_factorisations_bitwise_vs_reference(cases, fix):
    for name, kind, A, ic, b in cases:
        f = O.ilu_prepare(A, kind)
        assert np.array_equal(f, fix[f"{name}_factored"].reshape(f.shape)), name


def test_block_applies_bitwise_vs_reference(cases, fix):
    for name, kind, A, ic, b in cases:
        if kind == "ilu":
            continue
        f = O.ilu_prepare(A, kind)
        x = O.ilu_apply(A, f, b, kind, ic)
        assert np.array_equal(x, fix[f"{name}_apply"]), name


def test_scalar_apply_is_the_reference_minus_its_diagonal_term(cases, fix):
    """ilu.c:81 runs the backward loop from the diagonal slot: v -= values[dd]*b[i]
    before b[i] = v/values[dd].  Adding exactly that term to this oracle's
    sweep reproduces the reference's output bit for bit; without it the
    result is the exact triangular solve."""
    for name, kind, A, ic, b in cases:
        if kind != "ilu":
            continue
        f = O.ilu_prepare(A, kind).reshape(-1)
        rp, ci, ds = A.rowptr, A.colidx, A.diagslots
        x = np.array(b, np.float64)
        for i in range(A.bn):                        # forward (ilu.c:57-68)
            v = x[i]
            for j in range(rp[i], ds[i]):
                v -= f[j] * x[ci[j]]
            x[i] = v
        fwd = x.copy()
        for i in range(A.bn - 1, -1, -1):            # backward as in ilu.c:74-85
            v = x[i]
            for j in range(ds[i], rp[i + 1]):
                v -= f[j] * x[ci[j]]
            x[i] = v / f[ds[i]]
        assert np.array_equal(x, fix[f"{name}_apply"]), name
        ours = O.ilu_apply(A, f, b, "ilu")
        y = fwd.copy()
        for i in range(A.bn - 1, -1, -1):
            v = y[i]
            for j in range(ds[i] + 1, rp[i + 1]):
                v -= f[j] * y[ci[j]]
            y[i] = v / f[ds[i]]
        assert np.array_equal(ours, y), name
        assert not np.array_equal(ours, x)


def test_factorisation_reproduces_the_pattern(cases):
    """ILU(0): (L U)_ij = A_ij for every (i, j) of the pattern."""
    for name, kind, A, ic, b in cases:
        f = O.ilu_prepare(A, kind)
        Lm, Um = dense_factors(A, f, kind)
        LU = Lm @ Um
        blk = A.blk
        for i in range(A.bn):
            for p in range(A.rowptr[i], A.rowptr[i + 1]):
                c = A.colidx[p]
                got = LU[i * blk:(i + 1) * blk, c * blk:(c + 1) * blk]
                want = A.values[p].reshape(blk, blk)
                assert np.allclose(got, want, rtol=1e-11, atol=1e-11 * np.abs(want).max()), name


def test_applies_solve_with_the_factors(cases):
    for name, kind, A, ic, b in cases:
        f = O.ilu_prepare(A, kind)
        x = O.ilu_apply(A, f, b, kind, ic)
        Lm, Um = dense_factors(A, f, kind)
        r = Lm @ (Um @ x) - b
        assert np.linalg.norm(r) <= 1e-12 * np.linalg.norm(b), name


def test_preconditioned_solves_bitwise_vs_reference(cases, fix):
    """GMRES / BiCGSTAB with the block forms as preconditioner: iterations,
    status, residual history and solution equal the reference's bit for bit."""
    for name, kind, A, ic, b in cases:
        if kind == "ilu":
            continue
        for solver, kw in make_golden.ILU_SOLVES.items():
            if solver == "gmres":
                x, it, st, res, hist = O.gmres_ilu(A, b, kind, kw["m"], kw["maxiter"], kw["tol"],
                                                   icolors=ic)
            else:
                x, it, st, res, hist = O.bicgstab_ilu(A, b, kind, kw["maxiter"], kw["tol"],
                                                      icolors=ic)
            assert [it, st] == list(fix[f"{name}_{solver}_it_st"]), (name, solver)
            assert np.array_equal(hist[:it], fix[f"{name}_{solver}_hist"]), (name, solver)
            assert np.array_equal(x, fix[f"{name}_{solver}_x"]), (name, solver)


def test_colored_apply_equals_sequential_on_uncoupled_ranges(cases):
    for name, kind, A, ic, b in cases:
        if kind != "pbilu":
            continue
        f = O.ilu_prepare(A, kind)
        assert np.array_equal(O.ilu_apply(A, f, b, "pbilu", ic), O.ilu_apply(A, f, b, "bilu"))


@pytest.mark.skipif(not refkrylov.available(), reason="reference build absent")
def test_live_reference_on_another_pattern():
    A = make_golden.random_pattern_bsr(70, 4, 8, 411)
    b = np.random.default_rng(412).normal(size=A.bn * A.blk)
    vals, x = refkrylov.ilu_prepare_apply(A, "bilu", b)
    f = O.ilu_prepare(A, "bilu")
    assert np.array_equal(f, vals.reshape(f.shape))
    assert np.array_equal(O.ilu_apply(A, f, b, "bilu"), x)


@pytest.mark.skipif(not refkrylov.available(), reason="reference build absent")
@pytest.mark.parametrize("blk", [1, 3])
def test_live_reference_with_self_slots(blk):
    """Lower slots whose column is the row itself (the pivot block is then
    the row's own unfactored diagonal block, and the updates find the
    diagonal slot as the last match of that column)."""
    A = make_golden.with_self_slots(make_golden.random_pattern_bsr(80, blk, 6, 420 + blk), 5,
                                    421)
    b = np.random.default_rng(422).normal(size=A.bn * A.blk)
    kind = "bilu" if blk > 1 else "ilu"
    vals, x = refkrylov.ilu_prepare_apply(A, kind, b)
    f = O.ilu_prepare(A, kind)
    assert np.array_equal(f, vals.reshape(f.shape))
    if blk > 1:
        assert np.array_equal(O.ilu_apply(A, f, b, kind), x)

"""Pins the CPU oracle (oracle/lusgs_oracle.c) before it is trusted as the
checker of the GPU path:
  * against golden vectors written by the compiled reference
    (tests/golden/*.npz, from tests/golden/make_golden.py via the
    renumbering bridge) — expected bit for bit;
  * against the live reference in oracle/_ref where it is built here;
  * against closed forms and special cases the mathematics fixes.
"""
import hashlib
import os

import numpy as np
import pytest

import pyoracle as O
from ucfd_sparse_amd import mesh as M

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["hex8p", "hex6x5x4", "tet3p", "tet5p", "hex32p"]


def load_case(name):
    z = np.load(os.path.join(GOLD, f"lusgs_{name}.npz"))
    kind = str(z["kind"])
    params = tuple(int(v) for v in z["params"])
    seed = int(z["seed"])
    if kind == "hex":
        mesh = M.hex_mesh(*params[:3], seed=seed, permute=bool(params[3]))
    else:
        mesh = M.tet_mesh(params[0], seed=seed, permute=bool(params[1]))
    flow = M.flow_fields(mesh, seed=seed)
    return z, mesh, flow


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_flux_matches_reference_bitwise():
    z = np.load(os.path.join(GOLD, "flux_ref.npz"))
    for u, n, f in zip(z["u"], z["n"], z["f"]):
        assert np.array_equal(O.flux(u, n), f)


def test_flux_closed_forms():
    rng = np.random.default_rng(0)
    for _ in range(20):
        rho, p = rng.uniform(0.5, 2), rng.uniform(0.1, 3)
        vel = rng.uniform(-1, 1, 3)
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        E = p / 0.4 + 0.5 * rho * vel @ vel
        u = np.r_[rho, rho * vel, E]
        f = O.flux(u, n)
        un = vel @ n
        expect = np.r_[rho * un, rho * vel * un + n * p, (E + p) * un]
        np.testing.assert_allclose(f, expect, rtol=1e-13, atol=1e-13)
    # fluid at rest: only the pressure term survives
    u = np.r_[1.0, 0, 0, 0, 2.5]
    n = np.r_[0.6, 0.0, 0.8]
    np.testing.assert_allclose(O.flux(u, n), np.r_[0, 0.6, 0, 0.8, 0])
    # pressure floor: p -> PMIN and E rewritten to PMIN/(gamma-1) + ke
    u = np.r_[1.0, 0.5, 0, 0, 0.125 - 1e-3]
    f = O.flux(u, np.r_[1.0, 0, 0])
    pmin = 1e-13
    np.testing.assert_allclose(f, [0.5, 0.25 + pmin, 0, 0, (pmin / 0.4 + 0.125 + pmin) * 0.5],
                               rtol=1e-15)


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_fixture(name):
    z, mesh, flow = load_case(name)
    assert digest(mesh.nei_ele, mesh.fnorm_vol, mesh.vec_fnorm, flow.upts, flow.fspr,
                  flow.dt, flow.rhsb) == str(z["digest"]), "generator drifted"
    lcolor = z["lcolor"].astype(np.int32)
    icolor = np.argsort(lcolor, kind="stable").astype(np.int32)
    idx = z["sample"]
    upts = flow.upts.copy()
    iters = int(z["iterations"])
    for it in range(iters):
        diag, dstar, dq, upts = O.colored_pass(mesh.nei_ele, icolor, lcolor, mesh.fnorm_vol,
                                               mesh.vec_fnorm, flow.fspr, flow.dt, upts,
                                               flow.rhsb)
        if it == 0:
            assert np.array_equal(diag[idx], z["c_diag"])
            assert np.array_equal(dstar[:, idx], z["c_dstar"])
            assert np.array_equal(dq[:, idx], z["c_dq"])
            assert np.array_equal(upts[:, idx], z["c_upts1"])
    assert np.array_equal(upts[:, idx], z["c_uptsN"])
    if "s_dq" in z:
        diag, dstar, dq, u1 = O.serial_pass(mesh.nei_ele, mesh.fnorm_vol, mesh.vec_fnorm,
                                            flow.fspr, flow.dt, flow.upts, flow.rhsb)
        assert np.array_equal(dstar[:, idx], z["s_dstar"])
        assert np.array_equal(dq[:, idx], z["s_dq"])
        assert np.array_equal(u1[:, idx], z["s_upts1"])


def test_oracle_clamp_fixture():
    z = np.load(os.path.join(GOLD, "lusgs_clamp.npz"))
    mesh = M.hex_mesh(4, 4, 2, seed=3, permute=False)
    lcolor = z["lcolor"].astype(np.int32)
    icolor = np.argsort(lcolor, kind="stable").astype(np.int32)
    diag, dstar, dq, u1 = O.colored_pass(mesh.nei_ele, icolor, lcolor, mesh.fnorm_vol,
                                         mesh.vec_fnorm, z["fspr"], z["dt"], z["upts"], z["rhsb"])
    assert np.array_equal(dstar, z["c_dstar"])
    assert np.array_equal(dq, z["c_dq"])
    assert np.array_equal(u1, z["c_upts1"])


def test_first_color_and_all_self_mesh():
    mesh = M.tet_mesh(3, seed=9)
    flow = M.flow_fields(mesh, seed=9)
    from refbridge import greedy_coloring_py
    lcolor, icolor = greedy_coloring_py(mesh.nei_ele)
    diag, dstar, dq, _ = O.colored_pass(mesh.nei_ele, icolor, lcolor, mesh.fnorm_vol,
                                        mesh.vec_fnorm, flow.fspr, flow.dt, flow.upts, flow.rhsb)
    # diag closed form
    np.testing.assert_allclose(diag, 1 / flow.dt + 0.5 * (flow.fspr * mesh.fnorm_vol).sum(0),
                               rtol=1e-15)
    # colour 1 has no lower neighbour: dQ* = rhs / diag exactly
    c1 = lcolor == 1
    assert np.array_equal(dstar[:, c1], flow.rhsb[:, c1] / diag[c1])
    # the last colour has no upper neighbour: dQ = dQ*
    cl = lcolor == lcolor.max()
    assert np.array_equal(dq[:, cl], dstar[:, cl])
    # all faces on the boundary (self): the sweeps reduce to dQ = rhs/diag
    N = mesh.neles
    self_nei = np.tile(np.arange(N, dtype=np.int32), (4, 1))
    ones = np.ones(N, np.int32)
    diag, dstar, dq, _ = O.colored_pass(self_nei, np.arange(N, dtype=np.int32), ones,
                                        mesh.fnorm_vol, mesh.vec_fnorm, flow.fspr, flow.dt,
                                        flow.upts, flow.rhsb)
    assert np.array_equal(dq, flow.rhsb / diag)


def test_colored_equals_serial_on_color_ordering():
    """Renumbering bridge, checked inside the oracle: the coloured sweep equals
    the serial sweep run on the colour-ordered numbering."""
    mesh = M.hex_mesh(6, 5, 4, seed=4)
    flow = M.flow_fields(mesh, seed=4)
    from refbridge import greedy_coloring_py
    lcolor, icolor = greedy_coloring_py(mesh.nei_ele)
    N = mesh.neles
    rank = np.empty(N, np.int32)
    rank[icolor] = np.arange(N, dtype=np.int32)
    col = O.colored_pass(mesh.nei_ele, icolor, lcolor, mesh.fnorm_vol, mesh.vec_fnorm,
                         flow.fspr, flow.dt, flow.upts, flow.rhsb)

    def perm(a):
        out = np.empty_like(a)
        out[..., rank] = a
        return out
    nei_r = perm(rank[mesh.nei_ele])
    ser = O.serial_pass(nei_r, perm(mesh.fnorm_vol), perm(mesh.vec_fnorm), perm(flow.fspr),
                        perm(flow.dt), perm(flow.upts), perm(flow.rhsb))
    for a, b in zip(col, ser):
        assert np.array_equal(perm(a), b)


def test_oracle_matches_live_reference():
    import refbridge
    if not refbridge.available():
        pytest.skip("oracle/_ref not built here (needs the reference sources)")
    lib = refbridge.load()
    for seed, m in ((21, 4), (22, 6)):
        mesh = M.tet_mesh(m, seed=seed)
        flow = M.flow_fields(mesh, seed=seed)
        lcolor, icolor = refbridge.greedy_coloring_py(mesh.nei_ele)
        rank = np.empty(mesh.neles, np.int32)
        rank[icolor] = np.arange(mesh.neles, dtype=np.int32)
        ref = refbridge.lusgs_pass(lib, mesh.nei_ele, mesh.fnorm_vol, mesh.vec_fnorm, flow.fspr,
                                   flow.dt, flow.upts, flow.rhsb, rank=rank, iterations=2)
        upts = flow.upts.copy()
        for it in range(2):
            diag, dstar, dq, upts = O.colored_pass(mesh.nei_ele, icolor, lcolor, mesh.fnorm_vol,
                                                   mesh.vec_fnorm, flow.fspr, flow.dt, upts,
                                                   flow.rhsb)
            if it == 0:
                assert np.array_equal(diag, ref["diag"])
                assert np.array_equal(dstar, ref["dstar"])
                assert np.array_equal(dq, ref["dq"])
        assert np.array_equal(upts, ref["upts"][-1])

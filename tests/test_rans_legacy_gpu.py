"""GPU parity of the legacy RANS sweeps (parallel_rans_* / rans_parallel_*
and serial_rans_* / rans_serial_*, the {rans}_{serial,parallel}_{lower,
upper}_sweep family of doc/1_introduction.md:19-22) against the FlowSys
scalar LU-SGS oracle (oracle/flowsys_lusgs_oracle.c, pinned bit for bit to
the compiled reference's UCFDLUSGS_RANS* stages) through the renumbering
bridge: one slot per (cell, face) position, sides +1, face_area =
fnorm_vol, rcp_vol = 1, kappa = 1, a0 = dsrc = 0, tfspr per slot, cells in
colour-rank order.  The turbulence diagonal comes from parallel_pre_lusgs
with the turbulence spectral radius and factor = turb_factor, which is the
pack + RANS prepare of lusgs.c:23-36,127-137.  Bar: FP64 bit for bit
(asserted <= 1e-12 relative)."""
import sys

import numpy as np
import pytest

import pyoracle as O
from ucfd_sparse_amd import lib as L
from ucfd_sparse_amd import mesh as M
from ucfd_sparse_amd.layout import colored_layout

pytestmark = pytest.mark.gpu
RTOL = 1e-12
TURB_FACTOR = 1.5


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if L.device_count() == 0:
        pytest.fail("no GPU visible: the gpu tests must run on the MI355X")
    L.call("ucfd_set_stream", None)
    yield
    L.call("ucfd_set_config", 5, 5, 3)


def assert_close(a, b, what):
    e = np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
    say(f"  {what}: max rel diff {e:.3e} (bitwise {np.array_equal(a, b)})")
    assert e <= RTOL, f"{what}: {e}"


def quads(m, n, seed):
    h = M.hex_mesh(m, n, 1, seed=seed)
    return M.Mesh(np.ascontiguousarray(h.nei_ele[:4]), np.ascontiguousarray(h.fnorm_vol[:4]),
                  np.ascontiguousarray(h.vec_fnorm[:4, :2]), h.orig_id, h.key)


def with_turbulence(mesh, flow, nt, seed):
    """Positive turbulence states, N(0, 1e-3) right-hand sides and a
    turbulence spectral radius for rows NFVARS .. NFVARS+nt-1."""
    rng = np.random.default_rng(seed + 100)
    N, F = mesh.neles, mesh.nface
    nf = flow.upts.shape[0] - nt
    flow.upts[nf:] = rng.uniform(0.05, 1.0, (nt, N)) * flow.upts[0]
    flow.rhsb[nf:] = rng.normal(0.0, 1e-3, (nt, N))
    tfspr = rng.uniform(0.2, 1.2, (F, N))
    return flow, tfspr


def bridge_oracle(mesh, flow, tfspr, nfvars):
    """Turbulence diagonal, dQ* and dQ rows of one RANS lower + upper pass of
    the FlowSys oracle on the mesh's own numbering (the serial sweep; the
    coloured sweep on a colour-contiguous numbering)."""
    N, F = mesh.neles, mesh.nface
    nv, D = flow.upts.shape[0], mesh.ndims
    nei = np.ascontiguousarray(mesh.nei_ele)
    rowptr = np.arange(N + 1, dtype=np.int32) * F
    fs = M.FlowSysMesh(N, F * N, rowptr, np.ascontiguousarray(nei.T.reshape(-1), np.int32),
                       np.arange(F * N, dtype=np.int32), np.ones(F * N, np.int8),
                       np.ascontiguousarray(mesh.fnorm_vol.T.reshape(-1)),
                       np.ascontiguousarray(np.stack([mesh.vec_fnorm[:, d, :].T.reshape(-1)
                                                      for d in range(D)])),
                       np.ones(N), None)
    lam = np.ascontiguousarray(flow.fspr.T.reshape(-1))
    tlam = np.ascontiguousarray(tfspr.T.reshape(-1))
    el = {"cell_ids": np.arange(N, dtype=np.int32), "uptsb": flow.upts.copy(),
          "rhs": flow.rhsb.copy(), "dt": flow.dt.copy(), "dsrc": np.zeros_like(flow.upts),
          "vol": np.ones(N)}
    S = O.FlowSysLUSGSOracle(fs, lam, tlam, el, nv, nfvars)
    S.pack(TURB_FACTOR, 0.0)
    S.prepare(1.0, turb=True)
    tdiag = S.diag.reshape(nv, N)[nfvars].copy()
    S.sweep(False, 1.0, turb=True)
    dstar = S.du.reshape(nv, N)[nfvars:].copy()
    S.sweep(True, 1.0, turb=True)
    dq = S.du.reshape(nv, N)[nfvars:].copy()
    return tdiag, dstar, dq


def run_gpu(mesh, flow, tfspr, nfvars, colored, lay=None, two_array=False, device=True,
            declare=False):
    import torch
    N, F = mesh.neles, mesh.nface
    cv = (lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")) if device \
        else (lambda a: np.ascontiguousarray(a).copy())
    nei, fnv, vfn = cv(mesh.nei_ele), cv(mesh.fnorm_vol), cv(mesh.vec_fnorm)
    tsp, dt, upts = cv(tfspr), cv(flow.dt), cv(flow.upts)
    rhsb = cv(flow.rhsb)
    dub = cv(flow.rhsb) if not two_array else cv(np.zeros_like(flow.rhsb))
    tdiag = cv(np.zeros(N))
    host = (lambda a: a.cpu().numpy().copy()) if device else (lambda a: a.copy())
    L.call("parallel_pre_lusgs", N, F, TURB_FACTOR, fnv, dt, tdiag, tsp)
    if colored:
        ic, lc = cv(lay.icolor), cv(lay.lcolor)
        off = [int(x) for x in lay.off]
        if declare:
            L.call("ucfd_declare_colored_layout", N, ic, lc, None)
        try:
            for c in range(len(off) - 1):
                if two_array:
                    L.call("rans_parallel_lower_sweep", off[c], off[c + 1], N, F, nei, ic, lc, fnv,
                           vfn, upts, rhsb, dub, tdiag, tsp)
                else:
                    L.call("parallel_rans_lower_sweep", off[c], off[c + 1], N, F, nei, ic, lc, fnv,
                           vfn, upts, dub, tdiag, tsp)
            dstar = host(dub)[nfvars:]
            for c in reversed(range(len(off) - 1)):
                if two_array:
                    L.call("rans_parallel_upper_sweep", off[c], off[c + 1], N, F, nei, ic, lc, fnv,
                           vfn, upts, rhsb, dub, tdiag, tsp)
                else:
                    L.call("parallel_rans_upper_sweep", off[c], off[c + 1], N, F, nei, ic, lc, fnv,
                           vfn, upts, dub, tdiag, tsp)
        finally:
            if declare:
                L.call("ucfd_release_colored_layout", ic, lc)
    else:
        if two_array:
            L.call("rans_serial_lower_sweep", N, F, nei, fnv, vfn, upts, rhsb, dub, tdiag, tsp)
        else:
            L.call("serial_rans_lower_sweep", N, F, nei, fnv, vfn, upts, dub, tdiag, tsp)
        dstar = host(dub)[nfvars:]
        if two_array:
            L.call("rans_serial_upper_sweep", N, F, nei, fnv, vfn, upts, rhsb, dub, tdiag, tsp)
        else:
            L.call("serial_rans_upper_sweep", N, F, nei, fnv, vfn, upts, dub, tdiag, tsp)
    out = host(rhsb if two_array else dub)
    return host(tdiag), dstar, out


CASES = [("tet", 5, 1), ("tet", 5, 2), ("hex", 6, 2), ("quad", 9, 1)]


def make(kind, m, nt, seed):
    if kind == "tet":
        mesh = M.tet_mesh(m, seed=seed)
    elif kind == "hex":
        mesh = M.hex_mesh(m, m - 1, m - 2, seed=seed)
    else:
        mesh = quads(m, m - 2, seed)
    nfv = mesh.ndims + 2
    flow = M.flow_fields(mesh, seed=seed, nvars=nfv + nt)
    return mesh, flow, nfv


@pytest.mark.parametrize("kind, m, nt", CASES)
def test_colored_rans_sweeps_match_flowsys_oracle(kind, m, nt):
    seed = 31 + nt
    mesh, flow, nfv = make(kind, m, nt, seed)
    mesh.key = None
    lay = colored_layout(mesh, flow, use_key=False)
    m2, f2 = lay.mesh, lay.flow
    f2, tfspr = with_turbulence(m2, f2, nt, seed)
    say(f"{kind} {m2.neles} cells, nt={nt}, {lay.ncolors} colours")
    L.call("ucfd_set_config", nfv + nt, nfv, m2.ndims)
    od, os_, oq = bridge_oracle(m2, f2, tfspr, nfv)
    for device, two, declare in ((True, False, False), (True, True, False), (True, False, True),
                                 (True, True, True), (False, False, False)):
        tag = f"{'device' if device else 'host'} {'two-array' if two else 'in-place'}" \
              f"{' declared' if declare else ''}"
        td, ds, dq = run_gpu(m2, f2, tfspr, nfv, True, lay, two, device, declare)
        assert_close(td, od, f"{kind} nt={nt} tdiag ({tag})")
        assert_close(ds, os_, f"{kind} nt={nt} dQ* ({tag})")
        assert_close(dq[nfv:], oq, f"{kind} nt={nt} dQ ({tag})")
        if not two:
            # the flow rows of dub are not touched by the RANS sweeps
            assert np.array_equal(dq[:nfv], f2.rhsb[:nfv])


@pytest.mark.parametrize("kind, m, nt", CASES[:3])
def test_serial_rans_sweeps_match_flowsys_oracle(kind, m, nt):
    seed = 41 + nt
    mesh, flow, nfv = make(kind, m, nt, seed)
    flow, tfspr = with_turbulence(mesh, flow, nt, seed)
    L.call("ucfd_set_config", nfv + nt, nfv, mesh.ndims)
    od, os_, oq = bridge_oracle(mesh, flow, tfspr, nfv)
    for device, two in ((True, False), (True, True), (False, True)):
        td, ds, dq = run_gpu(mesh, flow, tfspr, nfv, False, None, two, device)
        tag = f"{'device' if device else 'host'} {'two-array' if two else 'in-place'}"
        assert_close(ds, os_, f"serial {kind} nt={nt} dQ* ({tag})")
        assert_close(dq[nfv:], oq, f"serial {kind} nt={nt} dQ ({tag})")


def test_rans_sweeps_refuse_a_configuration_without_turbulence_rows():
    mesh = M.tet_mesh(3, seed=3)
    flow = M.flow_fields(mesh, seed=3)
    N, F = mesh.neles, mesh.nface
    L.call("ucfd_set_config", 5, 5, 3)
    rc = L.load().serial_rans_lower_sweep(N, F, mesh.nei_ele.ctypes.data,
                                          mesh.fnorm_vol.ctypes.data, mesh.vec_fnorm.ctypes.data,
                                          flow.upts.ctypes.data, flow.rhsb.ctypes.data,
                                          flow.dt.ctypes.data, flow.fspr.ctypes.data)
    assert rc == 1

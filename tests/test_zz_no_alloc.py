"""ucfd_lusgs_residual_noalloc, the entry point that allocates nothing: its
residual and norm equal those of ucfd_lusgs_residual (the solver handle's
path) bit for bit, on two calls in a row, the second on a smaller mesh (the
fixed scratch keeps the first call's partial sums), for tets and hexes; and
it refuses what would need an allocation (host arrays) or more scratch than
it has."""
import ctypes
import sys

import numpy as np
import pytest

from ucfd_sparse_amd import lib as L
from ucfd_sparse_amd import mesh as M
from ucfd_sparse_amd.layout import colored_layout
from ucfd_sparse_amd.solver import LUSGSSolver

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if L.device_count() == 0:
        pytest.fail("no GPU visible: the gpu tests must run on the MI355X")
    L.call("ucfd_set_config", 5, 5, 3)
    L.call("ucfd_set_stream", None)
    yield


def layout(kind, m, seed, coloring):
    mesh = M.tet_mesh(m, seed=seed) if kind == "tet" else M.hex_mesh(m, m, m, seed=seed)
    flow = M.flow_fields(mesh, seed=seed)
    mesh.key = M.hilbert_keys(mesh, kind, (m,) if kind == "tet" else (m, m, m))
    return colored_layout(mesh, flow, coloring=coloring)


def compare(lay, seed):
    import torch
    dev = torch.device("cuda:0")

    def up(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    m, f = lay.mesh, lay.flow
    N, F = m.neles, m.nface
    nei, fnv, vfn, fsp, dt = (up(m.nei_ele), up(m.fnorm_vol), up(m.vec_fnorm), up(f.fspr),
                              up(f.dt))
    upts = up(np.ascontiguousarray(f.upts[:5]))
    b = up(f.rhsb[:5])
    rng = np.random.default_rng(seed)
    dq = up(rng.normal(0.0, 1e-3, size=(5, N)))
    S = LUSGSSolver(N, F, lay.off, nei, fnv, vfn, fsp, dt, upts)
    r_ref = torch.zeros_like(b)
    n_ref = S.residual(b, dq, r_ref)
    S.close()
    off = np.ascontiguousarray(lay.off, np.int32)
    r = torch.full_like(b, float("nan"))
    nrm = ctypes.c_double(-1.0)
    L.call("ucfd_lusgs_residual_noalloc", N, F, len(off) - 1, off.ctypes.data, nei, fnv, vfn,
           fsp, dt, upts, b, dq, r, ctypes.byref(nrm))
    a, e = r.cpu().numpy(), r_ref.cpu().numpy()
    err = np.abs(a - e).max() / np.abs(e).max()
    say(f"  N={N} F={F} colours={len(off) - 1}: r max rel diff {err:.3e} (bitwise "
        f"{np.array_equal(a, e)}), norm {nrm.value:.17g} vs {n_ref:.17g}")
    assert err <= RTOL
    assert abs(nrm.value - n_ref) <= RTOL * n_ref
    # the norm alone (r NULL) is the same
    nrm2 = ctypes.c_double(-1.0)
    L.call("ucfd_lusgs_residual_noalloc", N, F, len(off) - 1, off.ctypes.data, nei, fnv, vfn,
           fsp, dt, upts, b, dq, None, ctypes.byref(nrm2))
    assert nrm2.value == nrm.value


@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_noalloc_residual_equals_solver_residual_twice(kind):
    """Two calls in a row, the second on a smaller mesh."""
    big, small = (layout("tet", 24, 5, "greedy"), layout("tet", 9, 6, "greedy-bfs")) \
        if kind == "tet" else (layout("hex", 40, 7, "greedy"), layout("hex", 11, 8, "greedy"))
    say(f"{kind}: first call, {big.mesh.neles} cells")
    compare(big, 11)
    say(f"{kind}: second call, {small.mesh.neles} cells")
    compare(small, 12)


def test_noalloc_residual_refuses_host_arrays_and_oversize_meshes():
    lay = layout("tet", 4, 3, "greedy")
    m, f = lay.mesh, lay.flow
    N, F = m.neles, m.nface
    off = np.ascontiguousarray(lay.off, np.int32)
    b = np.ascontiguousarray(f.rhsb[:5])
    dq = np.zeros_like(b)
    rc = L.load().ucfd_lusgs_residual_noalloc(
        N, F, len(off) - 1, off.ctypes.data, m.nei_ele.ctypes.data, m.fnorm_vol.ctypes.data,
        m.vec_fnorm.ctypes.data, f.fspr.ctypes.data, f.dt.ctypes.data, f.upts.ctypes.data,
        b.ctypes.data, dq.ctypes.data, None, None)
    assert rc == 1
    big = 131072 * 256 + 1
    off2 = np.array([0, big], np.int32)
    rc = L.load().ucfd_lusgs_residual_noalloc(big, 4, 1, off2.ctypes.data, 8, 8, 8, 8, 8, 8, 8,
                                              8, None, None)
    assert rc == 1

"""GPU parity of the FP64 legacy LU-SGS API against the CPU oracle (itself
pinned bit for bit to the compiled reference) and the golden fixtures.

Bar: FP64 results within 1e-12 relative of the oracle (the kernels keep the
reference's operation order without FMA contraction, so they are expected
to agree bit for bit; the test reports the largest difference), integer
colouring bit-exact.  All calls go through the C-ABI (ctypes over
liblusgs.so), with host arrays (staged by the library) and with device
arrays (torch tensors, used in place)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import pyoracle as O
from refbridge import greedy_coloring_py
from ucfd_sparse_amd import lib as L
from ucfd_sparse_amd import layout as LY
from ucfd_sparse_amd import mesh as M

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(REPO, "tests", "golden")
RTOL = 1e-12


def _host(a):
    """A host copy of a numpy array or a device tensor."""
    return a.cpu().numpy().copy() if hasattr(a, "cpu") else np.array(a, copy=True)


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


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / scale


def assert_close(a, b, what):
    e = rel_err(a, b)
    say(f"  {what}: max rel diff {e:.3e} (bitwise equal: {np.array_equal(a, b)})")
    assert e <= RTOL, f"{what}: {e}"


def load_case(name):
    z = np.load(os.path.join(GOLD, f"lusgs_{name}.npz"))
    kind, params, seed = str(z["kind"]), [int(v) for v in z["params"]], int(z["seed"])
    mesh = (M.hex_mesh(*params[:3], seed=seed, permute=bool(params[3])) if kind == "hex"
            else M.tet_mesh(params[0], seed=seed, permute=bool(params[1])))
    flow = M.flow_fields(mesh, seed=seed)
    return z, mesh, flow


def colored_pass_lib(mesh, flow, lcolor, icolor, to_dev=None, two_array=False):
    """pre / lower per colour / upper per colour / update through the C-ABI."""
    N, F = mesh.neles, mesh.nface
    off = O.color_offsets(lcolor.cpu().numpy() if hasattr(lcolor, "cpu") else lcolor)
    cv = (lambda a: a) if to_dev is None else to_dev
    nei = cv(mesh.nei_ele)
    ic = icolor if hasattr(icolor, "data_ptr") else cv(icolor)
    lc = lcolor if hasattr(lcolor, "data_ptr") else cv(lcolor)
    fnv, vfn, fsp, dt = cv(mesh.fnorm_vol), cv(mesh.vec_fnorm), cv(flow.fspr), cv(flow.dt)
    upts = cv(flow.upts.copy())
    diag = cv(np.zeros(N))
    rhsb = cv(flow.rhsb.copy())
    dub = cv(flow.rhsb.copy()) if not two_array else cv(np.zeros_like(flow.rhsb))
    L.call("parallel_pre_lusgs", N, F, 1.0, fnv, dt, diag, fsp)
    nc = len(off) - 1
    for c in range(nc):
        if two_array:
            L.call("ns_parallel_lower_sweep", int(off[c]), int(off[c + 1]), N, F, nei, ic, lc,
                   fnv, vfn, upts, rhsb, dub, diag, fsp)
        else:
            L.call("parallel_ns_lower_sweep", int(off[c]), int(off[c + 1]), N, F, nei, ic, lc,
                   fnv, vfn, upts, dub, diag, fsp)
    dstar = _host(dub)
    for c in reversed(range(nc)):
        if two_array:
            L.call("ns_parallel_upper_sweep", int(off[c]), int(off[c + 1]), N, F, nei, ic, lc,
                   fnv, vfn, upts, rhsb, dub, diag, fsp)
        else:
            L.call("parallel_ns_upper_sweep", int(off[c]), int(off[c + 1]), N, F, nei, ic, lc,
                   fnv, vfn, upts, dub, diag, fsp)
    x = rhsb if two_array else dub
    L.call("lusgs_parallel_ns_update", N, upts, x)

    return _host(diag), dstar, _host(x), _host(upts)


def torch_dev():
    import torch

    def cv(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    return cv


@pytest.mark.parametrize("name", ["hex8p", "hex6x5x4", "tet3p", "tet5p", "hex32p"])
@pytest.mark.parametrize("where", ["host", "device"])
def test_colored_matches_oracle_and_fixture(name, where):
    say(f"case {name} ({where} arrays)")
    z, mesh, flow = load_case(name)
    lcolor = z["lcolor"].astype(np.int32)
    icolor = np.argsort(lcolor, kind="stable").astype(np.int32)
    to_dev = torch_dev() if where == "device" else None
    diag, dstar, dq, upts = colored_pass_lib(mesh, flow, lcolor, icolor, to_dev)
    od, os_, oq, ou = O.colored_pass(mesh.nei_ele, icolor, lcolor, mesh.fnorm_vol,
                                     mesh.vec_fnorm, flow.fspr, flow.dt, flow.upts, flow.rhsb)
    assert_close(diag, od, "diag")
    assert_close(dstar, os_, "dQ*")
    assert_close(dq, oq, "dQ")
    assert_close(upts, ou, "upts")
    idx = z["sample"]
    assert_close(dq[:, idx], z["c_dq"], "dQ vs reference fixture")


def isolate(mesh, cells):
    """The mesh with every face of the given cells, and the matching face of
    each of their neighbours, turned into a boundary self-face."""
    import dataclasses
    nei = mesh.nei_ele.copy()
    F = nei.shape[0]
    for i in cells:
        for f in range(F):
            j = int(nei[f, i])
            if j != i:
                nei[:, j][nei[:, j] == i] = j
            nei[f, i] = i
    return dataclasses.replace(mesh, nei_ele=nei)


@pytest.mark.parametrize("frac", [1.0, 0.15])
def test_boundary_only_cells(frac):
    """Cells whose faces are all boundary self-faces take no face term:
    dQ* = rhsb / diag and dQ = dQ* exactly; the rest of the mesh matches the
    oracle bit for bit (all cells isolated: one colour)."""
    say(f"tet mesh with {frac:.0%} of its cells isolated")
    mesh = M.tet_mesh(3, seed=91, permute=True)
    flow = M.flow_fields(mesh, seed=91)
    N = mesh.neles
    rng = np.random.default_rng(92)
    cells = np.arange(N) if frac == 1.0 else rng.choice(N, size=int(frac*N), replace=False)
    m2 = isolate(mesh, [int(c) for c in cells])
    alone = np.all(m2.nei_ele == np.arange(N)[None, :], axis=0)
    assert alone[cells].all()
    nc, lcolor, icolor, off = LY.greedy_coloring(m2.nei_ele)
    if frac == 1.0:
        assert nc == 1
    diag, dstar, dq, upts = colored_pass_lib(m2, flow, lcolor, icolor, torch_dev())
    od, os_, oq, ou = O.colored_pass(m2.nei_ele, icolor, lcolor, m2.fnorm_vol, m2.vec_fnorm,
                                     flow.fspr, flow.dt, flow.upts, flow.rhsb)
    for got, ref in ((diag, od), (dstar, os_), (dq, oq), (upts, ou)):
        assert np.array_equal(got, ref)
    assert np.array_equal(dstar[:, alone], flow.rhsb[:5, alone] / diag[alone])
    assert np.array_equal(dq[:, alone], dstar[:, alone])


@pytest.mark.parametrize("nvars,nfvars,ndims", [(4, 4, 2), (7, 5, 3), (12, 5, 3), (16, 5, 3)])
def test_update_every_row_count(nvars, nfvars, ndims):
    """upts += x over NVARS rows (lusgs.c:54-84): the kernel takes the cells in
    descending order and every row of a cell in one thread, in chunks of
    eight rows; host and device arrays, all four spellings, an odd cell
    count.  Elementwise addition: bit for bit against numpy."""
    import torch
    L.call("ucfd_set_config", nvars, nfvars, ndims)
    try:
        n = 1001
        rng = np.random.default_rng(nvars)
        for k, name in enumerate(("lusgs_parallel_ns_update", "parallel_ns_lusgs_update",
                                  "serial_ns_lusgs_update", "lusgs_serial_ns_update")):
            u0 = rng.normal(size=(nvars, n))
            x = rng.normal(size=(nvars, n))
            want = u0 + x
            if k % 2 == 0:
                u = u0.copy()
                L.call(name, n, u, x)
                got = u
            else:
                ud = torch.from_numpy(u0.copy()).cuda()
                L.call(name, n, ud, torch.from_numpy(x).cuda())
                L.call("ucfd_synchronize")
                got = ud.cpu().numpy()
            say(f"  {name}, nvars={nvars}: bitwise {np.array_equal(got, want)}")
            assert np.array_equal(got, want), name
    finally:
        L.call("ucfd_set_config", 5, 5, 3)


def test_two_array_form_equals_in_place_form():
    z, mesh, flow = load_case("tet5p")
    lcolor = z["lcolor"].astype(np.int32)
    icolor = np.argsort(lcolor, kind="stable").astype(np.int32)
    a = colored_pass_lib(mesh, flow, lcolor, icolor, torch_dev(), two_array=False)
    b = colored_pass_lib(mesh, flow, lcolor, icolor, torch_dev(), two_array=True)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("name", ["hex8p", "tet3p", "tet5p"])
def test_serial_matches_oracle(name):
    say(f"serial case {name}")
    z, mesh, flow = load_case(name)
    N, F = mesh.neles, mesh.nface
    cv = torch_dev()
    nei, fnv, vfn, fsp, dt = (cv(mesh.nei_ele), cv(mesh.fnorm_vol), cv(mesh.vec_fnorm),
                              cv(flow.fspr), cv(flow.dt))
    upts, dub, diag = cv(flow.upts.copy()), cv(flow.rhsb.copy()), cv(np.zeros(N))
    L.call("serial_pre_lusgs", N, F, 1.0, fnv, dt, diag, fsp)
    L.call("serial_ns_lower_sweep", N, F, nei, fnv, vfn, upts, dub, diag, fsp)
    dstar = dub.cpu().numpy().copy()
    L.call("serial_ns_upper_sweep", N, F, nei, fnv, vfn, upts, dub, diag, fsp)
    L.call("serial_ns_lusgs_update", N, upts, dub)
    od, os_, oq, ou = O.serial_pass(mesh.nei_ele, mesh.fnorm_vol, mesh.vec_fnorm, flow.fspr,
                                    flow.dt, flow.upts, flow.rhsb)
    assert_close(dstar, os_, "serial dQ*")
    assert_close(dub.cpu().numpy(), oq, "serial dQ")
    assert_close(upts.cpu().numpy(), ou, "serial upts")
    idx = z["sample"]
    assert_close(dub.cpu().numpy()[:, idx], z["s_dq"], "serial dQ vs reference fixture")
    # Python two-array spelling, host arrays
    upts_h, rhsb_h, dub_h, diag_h = flow.upts.copy(), flow.rhsb.copy(), np.zeros_like(flow.rhsb), \
        np.zeros(N)
    L.call("serial_pre_lusgs", N, F, 1.0, mesh.fnorm_vol, flow.dt, diag_h, flow.fspr)
    L.call("ns_serial_lower_sweep", N, F, mesh.nei_ele, mesh.fnorm_vol, mesh.vec_fnorm, upts_h,
           rhsb_h, dub_h, diag_h, flow.fspr)
    L.call("ns_serial_upper_sweep", N, F, mesh.nei_ele, mesh.fnorm_vol, mesh.vec_fnorm, upts_h,
           rhsb_h, dub_h, diag_h, flow.fspr)
    L.call("lusgs_serial_ns_update", N, upts_h, rhsb_h)
    assert_close(rhsb_h, oq, "serial two-array dQ")
    assert_close(upts_h, ou, "serial two-array upts")


def test_serial_schedule_follows_table_changes():
    """The serial sweeps keep their level schedule for the last neighbour
    table, keyed by a content hash: a table changed in place (same pointer)
    gets a new schedule, an unchanged one reuses it; both match the oracle."""
    say("serial sweeps: neighbour table changed in place")
    mesh = M.tet_mesh(4, seed=97, permute=True)
    flow = M.flow_fields(mesh, seed=97)
    N, F = mesh.neles, mesh.nface
    cv = torch_dev()
    m2 = isolate(mesh, [int(c) for c in np.random.default_rng(98).choice(N, 40, replace=False)])
    nei = cv(mesh.nei_ele)
    fnv, vfn, fsp, dt = cv(mesh.fnorm_vol), cv(mesh.vec_fnorm), cv(flow.fspr), cv(flow.dt)
    for m, label in ((mesh, "original"), (m2, "changed"), (m2, "unchanged")):
        nei.copy_(cv(m.nei_ele))
        upts, dub, diag = cv(flow.upts.copy()), cv(flow.rhsb.copy()), cv(np.zeros(N))
        L.call("serial_pre_lusgs", N, F, 1.0, fnv, dt, diag, fsp)
        L.call("serial_ns_lower_sweep", N, F, nei, fnv, vfn, upts, dub, diag, fsp)
        L.call("serial_ns_upper_sweep", N, F, nei, fnv, vfn, upts, dub, diag, fsp)
        od, os_, oq, ou = O.serial_pass(m.nei_ele, m.fnorm_vol, m.vec_fnorm, flow.fspr, flow.dt,
                                        flow.upts, flow.rhsb)
        say(f"  {label} table")
        assert np.array_equal(dub.cpu().numpy(), oq), label


def test_hex32_ten_iterations_match_reference_fixture():
    say("32^3 permuted hex, 10 iterations, rhsb restored each iteration")
    z, mesh, flow = load_case("hex32p")
    lcolor = z["lcolor"].astype(np.int32)
    icolor = np.argsort(lcolor, kind="stable").astype(np.int32)
    cv = torch_dev()
    u = flow.upts.copy()
    for _ in range(10):
        f2 = M.Flow(u, flow.fspr, flow.dt, flow.rhsb)
        _, _, _, u = colored_pass_lib(mesh, f2, lcolor, icolor, cv)
    assert_close(u[:, z["sample"]], z["c_uptsN"], "upts after 10 iterations vs reference")


def reorder(mesh, flow, lcolor):
    """Colour-contiguous renumbering with a Morton locality key."""
    lib = L.load()
    N, F = mesh.neles, mesh.nface
    perm = np.empty(N, np.int32)
    nc = int(lcolor.max())
    off = np.empty(nc + 1, np.int32)
    key = np.ascontiguousarray(mesh.key, np.uint64)
    lib.ucfd_color_reorder(N, lcolor.ctypes.data, key.ctypes.data, perm.ctypes.data,
                           off.ctypes.data)

    def p(a):
        out = np.empty_like(a)
        out[..., perm] = a
        return out
    nei = np.empty_like(mesh.nei_ele)
    lib.ucfd_permute_nei(N, F, perm.ctypes.data, mesh.nei_ele.ctypes.data, nei.ctypes.data)
    m2 = M.Mesh(nei, p(mesh.fnorm_vol), p(mesh.vec_fnorm), p(mesh.orig_id), p(mesh.key))
    f2 = M.Flow(p(flow.upts), p(flow.fspr), p(flow.dt), p(flow.rhsb))
    return m2, f2, p(lcolor), np.arange(N, dtype=np.int32), perm


def test_color_contiguous_layout_gives_identical_results():
    z, mesh, flow = load_case("tet5p")
    lcolor = z["lcolor"].astype(np.int32)
    icolor = np.argsort(lcolor, kind="stable").astype(np.int32)
    a = colored_pass_lib(mesh, flow, lcolor, icolor, torch_dev())
    m2, f2, lc2, ic2, perm = reorder(mesh, flow, lcolor)
    b = colored_pass_lib(m2, f2, lc2, ic2, torch_dev())
    for x, y in zip(a, b):
        assert np.array_equal(x, y[..., perm])


def test_declared_layout_fast_path_is_bitwise_identical():
    """ucfd_declare_colored_layout switches whole-colour calls to the
    index-range neighbour test; results must not change by a bit."""
    for name in ("tet5p", "hex32p"):
        z, mesh, flow = load_case(name)
        lcolor = z["lcolor"].astype(np.int32)
        icolor = np.argsort(lcolor, kind="stable").astype(np.int32)
        m2, f2, lc2, ic2, perm = reorder(mesh, flow, lcolor)
        cv = torch_dev()
        lc_d, ic_d = cv(lc2), cv(ic2)
        generic = colored_pass_lib(m2, f2, lc_d, ic_d, cv)
        nc = np.zeros(1, np.int32)
        L.call("ucfd_declare_colored_layout", m2.neles, ic_d, lc_d, nc)
        try:
            assert nc[0] == lcolor.max()
            fast = colored_pass_lib(m2, f2, lc_d, ic_d, cv)
            fast2 = colored_pass_lib(m2, f2, lc_d, ic_d, cv, two_array=True)
        finally:
            L.call("ucfd_release_colored_layout", ic_d, lc_d)
        for x, y, w in zip(generic, fast, fast2):
            assert np.array_equal(x, y) and np.array_equal(x, w)
        od, os_, oq, ou = O.colored_pass(m2.nei_ele, ic2, lc2, m2.fnorm_vol, m2.vec_fnorm,
                                         f2.fspr, f2.dt, f2.upts, f2.rhsb)
        assert_close(fast[2], oq, f"{name} declared-layout dQ")


def test_bad_neighbour_index_is_reported_not_faulted():
    z, mesh, flow = load_case("tet3p")
    lcolor = z["lcolor"].astype(np.int32)
    icolor = np.argsort(lcolor, kind="stable").astype(np.int32)
    nei = mesh.nei_ele.copy()
    nei[1, 5] = mesh.neles + 1000            # out of range: the kernel must skip and flag it
    N, F = mesh.neles, mesh.nface
    diag = np.zeros(N)
    L.call("parallel_pre_lusgs", N, F, 1.0, mesh.fnorm_vol, flow.dt, diag, flow.fspr)
    dub = flow.rhsb.copy()
    rc = L.load().parallel_ns_lower_sweep(0, N, N, F, nei.ctypes.data, icolor.ctypes.data,
                                          lcolor.ctypes.data, mesh.fnorm_vol.ctypes.data,
                                          mesh.vec_fnorm.ctypes.data, flow.upts.ctypes.data,
                                          dub.ctypes.data, diag.ctypes.data,
                                          flow.fspr.ctypes.data)
    assert rc == 3
    # the library keeps working after the reported error
    test_two_array_form_equals_in_place_form()


def _raw_pass_status(m, f, lc_d, ic_d, nei_d, cv):
    """Statuses of every call of one in-place legacy pass (device arrays)."""
    lib = L.load()
    N, F = m.neles, m.nface
    off = O.color_offsets(lc_d.cpu().numpy())
    fnv, vfn, fsp, dt, upts = (cv(m.fnorm_vol), cv(m.vec_fnorm), cv(f.fspr), cv(f.dt),
                               cv(f.upts.copy()))
    diag, dub = cv(np.zeros(N)), cv(f.rhsb.copy())
    L.call("parallel_pre_lusgs", N, F, 1.0, fnv, dt, diag, fsp)
    args = lambda c: (int(off[c]), int(off[c + 1]), N, F, nei_d.data_ptr(),  # noqa: E731
                      ic_d.data_ptr(), lc_d.data_ptr(), fnv.data_ptr(), vfn.data_ptr(),
                      upts.data_ptr(), dub.data_ptr(), diag.data_ptr(), fsp.data_ptr())
    nc = len(off) - 1
    low = [lib.parallel_ns_lower_sweep(*args(c)) for c in range(nc)]
    upp = [lib.parallel_ns_upper_sweep(*args(c)) for c in reversed(range(nc))]
    return low, upp


def test_declared_layout_reports_changed_arrays():
    """A declared layout is found by its pointers; if lcolor changes under
    the same address (as after a free and reuse), the range-mode calls
    sample it and fail with status 5 instead of sweeping with the wrong
    neighbour test.  Re-declaring restores the fast path."""
    z, mesh, flow = load_case("tet5p")
    lcolor = z["lcolor"].astype(np.int32)
    m2, f2, lc2, ic2, perm = reorder(mesh, flow, lcolor)
    cv = torch_dev()
    lc_d, ic_d, nei_d = cv(lc2), cv(ic2), cv(m2.nei_ele)
    L.call("ucfd_declare_colored_layout", m2.neles, ic_d, lc_d, None)
    try:
        low, upp = _raw_pass_status(m2, f2, lc_d, ic_d, nei_d, cv)
        assert low == [0] * len(low) and upp == [0] * len(upp)
        # another colouring written into the same buffer
        lc_d.copy_(cv(lc2[::-1].copy()))
        low, upp = _raw_pass_status(m2, f2, lc_d, ic_d, nei_d, cv)
        say(f"  statuses after the arrays changed: lower {low} upper {upp}")
        assert 5 in low + upp
        lc_d.copy_(cv(lc2))
        L.call("ucfd_declare_colored_layout", m2.neles, ic_d, lc_d, None)
        low, upp = _raw_pass_status(m2, f2, lc_d, ic_d, nei_d, cv)
        assert low == [0] * len(low) and upp == [0] * len(upp)
    finally:
        L.call("ucfd_release_colored_layout", ic_d, lc_d)
    test_two_array_form_equals_in_place_form()


def test_bad_index_in_face_free_range_is_reported_by_the_other_sweep():
    """On a declared layout the first colour of the lower sweep takes no face
    and reads no neighbour index (k_sweep_none): an out-of-range entry of a
    colour-1 cell is not seen by that call, and the upper sweep of colour 1,
    which reads every index of its cells, reports it (status 3)."""
    z, mesh, flow = load_case("tet5p")
    lcolor = z["lcolor"].astype(np.int32)
    m2, f2, lc2, ic2, perm = reorder(mesh, flow, lcolor)
    nei = m2.nei_ele.copy()
    nei[2, 3] = m2.neles + 7                 # cell 3 lies in colour 1 (the first range)
    assert lc2[3] == 1
    cv = torch_dev()
    lc_d, ic_d, nei_d = cv(lc2), cv(ic2), cv(nei)
    L.call("ucfd_declare_colored_layout", m2.neles, ic_d, lc_d, None)
    try:
        low, upp = _raw_pass_status(m2, f2, lc_d, ic_d, nei_d, cv)
    finally:
        L.call("ucfd_release_colored_layout", ic_d, lc_d)
    say(f"  lower statuses {low}, upper statuses (last colour first) {upp}")
    assert low[0] == 0
    assert upp[-1] == 3
    test_two_array_form_equals_in_place_form()


def test_async_stream_mode_matches_sync():
    import torch
    z, mesh, flow = load_case("tet5p")
    lcolor = z["lcolor"].astype(np.int32)
    icolor = np.argsort(lcolor, kind="stable").astype(np.int32)
    a = colored_pass_lib(mesh, flow, lcolor, icolor, torch_dev())
    s = torch.cuda.Stream()
    L.call("ucfd_set_stream", s.cuda_stream)
    try:
        cv = torch_dev()
        N, F = mesh.neles, mesh.nface
        off = O.color_offsets(lcolor)
        nei, ic, lc = cv(mesh.nei_ele), cv(icolor), cv(lcolor)
        fnv, vfn, fsp, dt = cv(mesh.fnorm_vol), cv(mesh.vec_fnorm), cv(flow.fspr), cv(flow.dt)
        upts, dub, diag = cv(flow.upts.copy()), cv(flow.rhsb.copy()), cv(np.zeros(N))
        torch.cuda.synchronize()
        L.call("parallel_pre_lusgs", N, F, 1.0, fnv, dt, diag, fsp)
        for c in range(len(off) - 1):
            L.call("parallel_ns_lower_sweep", int(off[c]), int(off[c + 1]), N, F, nei, ic, lc,
                   fnv, vfn, upts, dub, diag, fsp)
        for c in reversed(range(len(off) - 1)):
            L.call("parallel_ns_upper_sweep", int(off[c]), int(off[c + 1]), N, F, nei, ic, lc,
                   fnv, vfn, upts, dub, diag, fsp)
        L.call("lusgs_parallel_ns_update", N, upts, dub)
        L.call("ucfd_synchronize")
    finally:
        L.call("ucfd_set_stream", None)
    assert np.array_equal(upts.cpu().numpy(), a[3])
    assert np.array_equal(dub.cpu().numpy(), a[2])


def test_determinism():
    z, mesh, flow = load_case("hex8p")
    lcolor = z["lcolor"].astype(np.int32)
    icolor = np.argsort(lcolor, kind="stable").astype(np.int32)
    a = colored_pass_lib(mesh, flow, lcolor, icolor, torch_dev())
    b = colored_pass_lib(mesh, flow, lcolor, icolor, torch_dev())
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_two_dimensional_quads():
    """NDIMS = 2 (nfvars = 4) through ucfd_set_config, on a quad grid."""
    h = M.hex_mesh(9, 7, 1, seed=6)
    N = h.neles
    m2 = M.Mesh(np.ascontiguousarray(h.nei_ele[:4]), np.ascontiguousarray(h.fnorm_vol[:4]),
                np.ascontiguousarray(h.vec_fnorm[:4, :2]), h.orig_id, h.key)
    flow = M.flow_fields(m2, seed=6)
    lcolor, icolor = greedy_coloring_py(m2.nei_ele)
    L.call("ucfd_set_config", 4, 4, 2)
    try:
        a = colored_pass_lib(m2, flow, lcolor, icolor, torch_dev())
    finally:
        L.call("ucfd_set_config", 5, 5, 3)
    od, os_, oq, ou = O.colored_pass(m2.nei_ele, icolor, lcolor, m2.fnorm_vol, m2.vec_fnorm,
                                     flow.fspr, flow.dt, flow.upts, flow.rhsb)
    assert_close(a[2], oq, "2-D dQ")
    assert_close(a[3], ou, "2-D upts")


def test_two_dimensional_declared_layout_bitwise():
    """NDIMS = 2 through a declared colour-contiguous layout: the range-mode
    sweeps (the all-taken variant, and the ranges with no face in the sweep
    direction, which skip the neighbour indices) equal the generic colour
    test bit for bit, in both array forms, and follow the oracle."""
    h = M.hex_mesh(9, 7, 1, seed=6)
    m = M.Mesh(np.ascontiguousarray(h.nei_ele[:4]), np.ascontiguousarray(h.fnorm_vol[:4]),
               np.ascontiguousarray(h.vec_fnorm[:4, :2]), h.orig_id, h.key)
    flow = M.flow_fields(m, seed=6)
    lcolor, _ = greedy_coloring_py(m.nei_ele)
    m2, f2, lc2, ic2, _ = reorder(m, flow, lcolor)
    say(f"2-D declared layout: {m2.neles} quads, {int(lc2.max())} colours")
    L.call("ucfd_set_config", 4, 4, 2)
    try:
        cv = torch_dev()
        lc_d, ic_d = cv(lc2), cv(ic2)
        generic = colored_pass_lib(m2, f2, lc_d, ic_d, cv)
        L.call("ucfd_declare_colored_layout", m2.neles, ic_d, lc_d, None)
        try:
            fast = colored_pass_lib(m2, f2, lc_d, ic_d, cv)
            fast2 = colored_pass_lib(m2, f2, lc_d, ic_d, cv, two_array=True)
        finally:
            L.call("ucfd_release_colored_layout", ic_d, lc_d)
    finally:
        L.call("ucfd_set_config", 5, 5, 3)
    for x, y, w in zip(generic, fast, fast2):
        assert np.array_equal(x, y) and np.array_equal(x, w)
    od, os_, oq, ou = O.colored_pass(m2.nei_ele, ic2, lc2, m2.fnorm_vol, m2.vec_fnorm,
                                     f2.fspr, f2.dt, f2.upts, f2.rhsb)
    assert_close(fast[2], oq, "2-D declared-layout dQ")
    assert_close(fast[3], ou, "2-D declared-layout upts")


def test_c_abi_program_matches_oracle():
    exe = os.path.join(REPO, "tests", "abi", "legacy_calls.x")
    if not os.path.exists(exe):
        pytest.fail("tests/abi/legacy_calls.x not built (run __graft_entry__.build())")
    out = os.path.join(REPO, "gpurun_out", "legacy_calls.bin")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    env = dict(os.environ, LD_LIBRARY_PATH=os.path.join(REPO, "ucfd_sparse_amd"))
    r = subprocess.run([exe, out, "6", "5", "4"], capture_output=True, text=True, timeout=120,
                       env=env)
    say(r.stdout.strip())
    assert r.returncode == 0, r.stderr
    raw = open(out, "rb").read()
    m, n, l, N = np.frombuffer(raw[:16], np.int32)
    pos = 16

    def take(dtype, count):
        nonlocal pos
        a = np.frombuffer(raw[pos:pos + count * np.dtype(dtype).itemsize], dtype)
        pos += count * np.dtype(dtype).itemsize
        return a
    nei = take(np.int32, 6 * N).reshape(6, N)
    icolor = take(np.int32, N)
    lcolor = take(np.int32, N)
    fnv = take(np.float64, 6 * N).reshape(6, N)
    vfn = take(np.float64, 18 * N).reshape(6, 3, N)
    fspr = take(np.float64, 6 * N).reshape(6, N)
    dt = take(np.float64, N)
    upts0 = take(np.float64, 5 * N).reshape(5, N)
    rhs0 = take(np.float64, 5 * N).reshape(5, N)
    od, os_, oq, ou = O.colored_pass(nei, icolor, lcolor, fnv, vfn, fspr, dt, upts0, rhs0)
    assert_close(take(np.float64, N), od, "C caller diag")
    assert_close(take(np.float64, 5 * N).reshape(5, N), os_, "C caller dQ*")
    assert_close(take(np.float64, 5 * N).reshape(5, N), oq, "C caller dQ")
    assert_close(take(np.float64, 5 * N).reshape(5, N), ou, "C caller upts")
    take(np.float64, 5 * N)
    assert_close(take(np.float64, 5 * N).reshape(5, N), oq, "C caller two-array dQ")
    assert_close(take(np.float64, 5 * N).reshape(5, N), ou, "C caller two-array upts")
    sd, ss, sq, su = O.serial_pass(nei, fnv, vfn, fspr, dt, upts0, rhs0)
    assert_close(take(np.float64, 5 * N).reshape(5, N), sq, "C caller serial dQ")
    assert_close(take(np.float64, 5 * N).reshape(5, N), su, "C caller serial upts")


@pytest.fixture(scope="module")
def tet96():
    say("96^3x6 tets (5.3M cells): generating the mesh, fields and Hilbert keys")
    mesh = M.tet_mesh(96, seed=1)
    flow = M.flow_fields(mesh, seed=1)
    mesh.key = M.hilbert_keys(mesh, "tet", (96,))
    return mesh, flow


@pytest.mark.parametrize("coloring", ["greedy", "greedy-bfs"])
def test_tet96_5m_cells_fp64_matches_oracle(tet96, coloring):
    """Full benchmark size (96^3 x 6 tets = 5,308,416 cells), permuted, in the
    layout the benchmark times: coloured by the a'1 first-fit routine
    ("greedy", 5 colours) or by first-fit in breadth-first order
    ("greedy-bfs", the benchmark's 2 colours), renumbered colour-contiguously
    in Hilbert order and declared, so the range-mode kernels run
    (k_sweep_none for the empty ranges, the all-taken variant for the last
    lower / first upper colour).  Both array forms; one pass against the
    OpenMP oracle."""
    import torch
    mesh, flow = tet96
    say(f"96^3x6 tets (5.3M cells), {coloring}: colour-contiguous Hilbert layout")
    lay = LY.colored_layout(mesh, flow, coloring=coloring)
    m2, f2, N = lay.mesh, lay.flow, lay.mesh.neles
    say(f"96^3x6 tets, {coloring}: {lay.ncolors} colours {np.diff(lay.off).tolist()}")
    cv = torch_dev()
    lc_d, ic_d = cv(lay.lcolor), cv(lay.icolor)
    L.call("ucfd_declare_colored_layout", N, ic_d, lc_d, None)
    try:
        say(f"96^3x6 tets, {coloring}: GPU pass, in-place C form")
        diag, dstar, dq, upts = colored_pass_lib(m2, f2, lc_d, ic_d, cv)
        torch.cuda.empty_cache()
        say(f"96^3x6 tets, {coloring}: GPU pass, two-array form (as benchmarked)")
        _, dstar2, dq2, upts2 = colored_pass_lib(m2, f2, lc_d, ic_d, cv, two_array=True)
    finally:
        L.call("ucfd_release_colored_layout", ic_d, lc_d)
    del lc_d, ic_d
    torch.cuda.empty_cache()
    say(f"96^3x6 tets, {coloring}: oracle pass")
    od, os_, oq, ou = O.colored_pass(m2.nei_ele, lay.icolor, lay.lcolor, m2.fnorm_vol,
                                     m2.vec_fnorm, f2.fspr, f2.dt, f2.upts, f2.rhsb)
    assert_close(diag, od, f"tet96 {coloring} diag")
    assert_close(dstar, os_, f"tet96 {coloring} dQ*")
    assert_close(dq, oq, f"tet96 {coloring} dQ")
    assert_close(upts, ou, f"tet96 {coloring} upts")
    assert_close(dstar2, os_, f"tet96 {coloring} two-array dQ*")
    assert_close(dq2, oq, f"tet96 {coloring} two-array dQ")
    assert_close(upts2, ou, f"tet96 {coloring} two-array upts")


def test_pressure_clamp_fixture_on_gpu():
    """The legacy kernels on cells at the pressure floor: the reference's
    outputs for a case where PMIN clamps (flux.c:50-56) the states of 18 of
    32 cells, tests/golden/lusgs_clamp.npz (written by the compiled reference
    through the renumbering bridge on an unpermuted 4x4x2 hex grid),
    reproduced by the GPU pass through the C-ABI, host and device arrays."""
    z = np.load(os.path.join(GOLD, "lusgs_clamp.npz"))
    mesh = M.hex_mesh(4, 4, 2, seed=3, permute=False)
    flow = M.Flow(z["upts"], z["fspr"], z["dt"], z["rhsb"])
    u = z["upts"]
    p = 0.4 * (u[4] - 0.5 * (u[1] ** 2 + u[2] ** 2 + u[3] ** 2) / u[0])
    say(f"clamp fixture: {(p < 1e-13).sum()} of {p.size} cells below PMIN")
    assert (p < 1e-13).sum() > 0
    lcolor = z["lcolor"].astype(np.int32)
    icolor = np.argsort(lcolor, kind="stable").astype(np.int32)
    for to_dev in (None, torch_dev()):
        diag, dstar, dq, upts = colored_pass_lib(mesh, flow, lcolor, icolor, to_dev)
        where = "host" if to_dev is None else "device"
        assert_close(diag, z["c_diag"], f"clamp diag ({where})")
        assert_close(dstar, z["c_dstar"], f"clamp dQ* ({where})")
        assert_close(dq, z["c_dq"], f"clamp dQ ({where})")
        assert_close(upts, z["c_upts1"], f"clamp upts ({where})")

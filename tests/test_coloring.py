"""Greedy colouring and colour-contiguous renumbering (host routines of
liblusgs.so; no GPU needed).  Pinned by the red-black known answer of
examples/utils/pbutils.py:39-56 on unpermuted hex grids, by an independent
Python first-fit, by the colourings stored in the reference fixtures, and by
validity / bound / round-trip invariants."""
import os

import numpy as np
import pytest

from refbridge import greedy_coloring_py
from ucfd_sparse_amd import lib as L
from ucfd_sparse_amd import mesh as M

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def color(nei):
    nei = np.ascontiguousarray(nei, np.int32)
    F, N = nei.shape
    lcolor = np.empty(N, np.int32)
    icolor = np.empty(N, np.int32)
    off = np.empty(2 * F + 4, np.int32)
    nc = L.load().ucfd_greedy_coloring(N, F, nei.ctypes.data, icolor.ctypes.data,
                                       lcolor.ctypes.data, off.ctypes.data)
    assert nc > 0
    return nc, lcolor, icolor, off[:nc + 1]


def pbutils_red_black(m, n, l):
    """Known answer restating examples/utils/pbutils.py:55-72 (make_coloring):
    even (x+y+z) cells first, then odd, each in ascending id."""
    i = np.arange(m * n * l)
    x, y, z = i % m, (i // m) % n, i // (m * n)
    even = ((x + y + z) & 1) == 0
    return np.concatenate((i[even], i[~even])).astype(np.int32), np.where(even, 1, 2)


@pytest.mark.parametrize("dims", [(4, 4, 4), (5, 3, 2), (6, 5, 4), (32, 32, 32)])
def test_red_black_known_answer(dims):
    mesh = M.hex_mesh(*dims, seed=0, permute=False)
    nc, lcolor, icolor, off = color(mesh.nei_ele)
    ic_ref, lc_ref = pbutils_red_black(*dims)
    assert nc == 2
    assert np.array_equal(lcolor, lc_ref)
    assert np.array_equal(icolor, ic_ref)


@pytest.mark.parametrize("kind,arg,seed", [("hex", (8, 8, 8), 1), ("hex", (9, 7, 5), 2),
                                           ("tet", 4, 3), ("tet", 7, 5)])
def test_greedy_matches_independent_statement(kind, arg, seed):
    mesh = M.hex_mesh(*arg, seed=seed) if kind == "hex" else M.tet_mesh(arg, seed=seed)
    nc, lcolor, icolor, off = color(mesh.nei_ele)
    lc_py, ic_py = greedy_coloring_py(mesh.nei_ele)
    assert np.array_equal(lcolor, lc_py)
    assert np.array_equal(icolor, ic_py)
    assert nc <= mesh.nface + 1
    assert L.load().ucfd_coloring_conflicts(mesh.neles, mesh.nface, mesh.nei_ele.ctypes.data,
                                            lcolor.ctypes.data) == 0
    assert np.array_equal(np.diff(off), np.bincount(lcolor)[1:])


@pytest.mark.parametrize("name", ["hex8p", "tet3p", "tet5p", "hex32p"])
def test_greedy_matches_fixture_coloring(name):
    z = np.load(os.path.join(GOLD, f"lusgs_{name}.npz"))
    kind, params, seed = str(z["kind"]), [int(v) for v in z["params"]], int(z["seed"])
    mesh = (M.hex_mesh(*params[:3], seed=seed, permute=bool(params[3])) if kind == "hex"
            else M.tet_mesh(params[0], seed=seed, permute=bool(params[1])))
    nc, lcolor, icolor, off = color(mesh.nei_ele)
    assert np.array_equal(lcolor, z["lcolor"].astype(np.int32))


def test_hex32_permuted_coloring_has_six_colors():
    mesh = M.hex_mesh(32, 32, 32, seed=0)
    nc, lcolor, icolor, off = color(mesh.nei_ele)
    assert nc == 6
    assert np.array_equal(icolor, np.argsort(lcolor, kind="stable"))


def test_greedy_is_deterministic_and_valid_on_nonsymmetric_input():
    rng = np.random.default_rng(5)
    N, F = 300, 4
    nei = rng.integers(0, N, size=(F, N)).astype(np.int32)     # arbitrary, not symmetric
    a = color(nei)
    b = color(nei)
    assert all(np.array_equal(x, y) for x, y in zip(a[1:], b[1:]))
    assert L.load().ucfd_coloring_conflicts(N, F, nei.ctypes.data, a[1].ctypes.data) == 0


def test_reorder_round_trip_and_contiguity():
    mesh = M.tet_mesh(5, seed=8)
    N, F = mesh.neles, mesh.nface
    nc, lcolor, icolor, off = color(mesh.nei_ele)
    lib = L.load()
    perm = np.empty(N, np.int32)
    off2 = np.empty(nc + 1, np.int32)
    key = np.ascontiguousarray(mesh.key, np.uint64)
    assert lib.ucfd_color_reorder(N, lcolor.ctypes.data, key.ctypes.data, perm.ctypes.data,
                                  off2.ctypes.data) == nc
    assert np.array_equal(np.sort(perm), np.arange(N))
    assert np.array_equal(off, off2)
    lc_new = np.empty(N, np.int32)
    lib.ucfd_permute_rows(N, 1, 4, perm.ctypes.data, lcolor.ctypes.data, lc_new.ctypes.data)
    assert np.all(np.diff(lc_new) >= 0)                        # colours contiguous, ascending
    nei_new = np.empty_like(mesh.nei_ele)
    lib.ucfd_permute_nei(N, F, perm.ctypes.data, mesh.nei_ele.ctypes.data, nei_new.ctypes.data)
    # unpermute: old = inv[new]
    inv = np.empty(N, np.int32)
    inv[perm] = np.arange(N, dtype=np.int32)
    assert np.array_equal(inv[nei_new[:, perm]], mesh.nei_ele)
    vfn_new = np.empty_like(mesh.vec_fnorm)
    lib.ucfd_permute_rows(N, F * 3, 8, perm.ctypes.data, mesh.vec_fnorm.ctypes.data,
                          vfn_new.ctypes.data)
    assert np.array_equal(vfn_new[..., perm], mesh.vec_fnorm)
    # within a colour the locality key is ascending
    k_new = np.empty(N, np.uint64)
    k_new[perm] = key
    for c in range(nc):
        assert np.all(np.diff(k_new[off[c]:off[c + 1]].astype(np.int64)) >= 0)
    # the renumbered colouring is still valid
    assert lib.ucfd_coloring_conflicts(N, F, nei_new.ctypes.data, lc_new.ctypes.data) == 0


@pytest.mark.slow
def test_benchmark_mesh_colorings_fixture():
    path = os.path.join(GOLD, "colorings.npz")
    if not os.path.exists(path):
        pytest.skip("colorings.npz not generated")
    import hashlib
    z = np.load(path)
    for name, m, seed in (("tet55p", 55, 4), ("tet70p", 70, 3), ("tet96p", 96, 1)):
        mesh = M.tet_mesh(m, seed=seed)
        nc, lcolor, icolor, off = color(mesh.nei_ele)
        assert np.array_equal(np.bincount(lcolor)[1:], z[name + "_counts"])
        assert hashlib.sha256(lcolor.tobytes()).hexdigest() == str(z[name + "_digest"])


def test_declared_layout_validation_on_host_arrays():
    lib = L.load()
    N = 50
    lc = np.repeat(np.arange(1, 6, dtype=np.int32), 10)
    ic = np.arange(N, dtype=np.int32)
    nc = np.zeros(1, np.int32)
    assert lib.ucfd_declare_colored_layout(N, ic.ctypes.data, lc.ctypes.data, nc.ctypes.data) == 0
    assert nc[0] == 5
    assert lib.ucfd_release_colored_layout(ic.ctypes.data, lc.ctypes.data) == 0
    bad_ic = ic.copy()
    bad_ic[[3, 4]] = bad_ic[[4, 3]]
    assert lib.ucfd_declare_colored_layout(N, bad_ic.ctypes.data, lc.ctypes.data, None) != 0
    bad_lc = lc.copy()
    bad_lc[7] = 3
    assert lib.ucfd_declare_colored_layout(N, ic.ctypes.data, bad_lc.ctypes.data, None) != 0


def test_balanced_coloring_is_valid_balanced_and_deterministic():
    from ucfd_sparse_amd import lib as L
    from ucfd_sparse_amd.layout import balanced_coloring
    mesh = M.tet_mesh(12, seed=5, permute=True)
    nei = np.ascontiguousarray(mesh.nei_ele)
    for target in (0, 5, 7):
        nc, lcolor, icolor, off = balanced_coloring(nei, target)
        assert nc == max(target, 5)                      # tets: 4 distinct neighbours
        assert L.load().ucfd_coloring_conflicts(mesh.neles, 4, nei.ctypes.data,
                                                lcolor.ctypes.data) == 0
        sizes = np.diff(off)
        assert sizes.max() - sizes.min() <= max(2, sizes.max() // 100)
        assert np.array_equal(np.sort(icolor), np.arange(mesh.neles))
        assert np.all(np.diff(lcolor[icolor]) >= 0)
        nc2, lcolor2, _, _ = balanced_coloring(nei, target)
        assert np.array_equal(lcolor, lcolor2)


def test_ordered_greedy_coloring():
    from ucfd_sparse_amd.layout import greedy_coloring, greedy_coloring_ordered
    mesh = M.tet_mesh(10, seed=9, permute=True)
    nei = np.ascontiguousarray(mesh.nei_ele)
    nc, lc, ic, off = greedy_coloring(nei)
    nc2, lc2, ic2, off2 = greedy_coloring_ordered(nei, np.arange(mesh.neles, dtype=np.int32))
    assert nc == nc2 and np.array_equal(lc, lc2) and np.array_equal(ic, ic2)
    order = np.argsort(M.hilbert_keys(mesh, "tet", (10,)), kind="stable").astype(np.int32)
    nc3, lc3, ic3, off3 = greedy_coloring_ordered(nei, order)
    assert nc3 <= 5
    assert L.load().ucfd_coloring_conflicts(mesh.neles, 4, nei.ctypes.data, lc3.ctypes.data) == 0
    # visiting along the curve packs most cells into the first two colours
    assert (off3[2] - off3[0]) > 0.75 * mesh.neles
    bad = order.copy()
    bad[1] = bad[0]
    with pytest.raises(ValueError):
        greedy_coloring_ordered(nei, bad)


def test_reduce_colors_removes_the_top_colour():
    from ucfd_sparse_amd.layout import greedy_coloring_ordered, reduce_colors
    mesh = M.tet_mesh(14, seed=2, permute=True)
    nei = np.ascontiguousarray(mesh.nei_ele)
    order = np.argsort(M.hilbert_keys(mesh, "tet", (14,)), kind="stable").astype(np.int32)
    nc, lc, ic, off = greedy_coloring_ordered(nei, order)
    nc2, lc2, ic2, off2 = reduce_colors(nei, lc)
    assert nc2 <= nc
    assert L.load().ucfd_coloring_conflicts(mesh.neles, 4, nei.ctypes.data, lc2.ctypes.data) == 0
    assert np.array_equal(np.diff(off2), np.bincount(lc2)[1:])
    assert np.array_equal(np.sort(ic2), np.arange(mesh.neles))
    # cells that did not have the top colour keep theirs unless a chain moved them
    changed = np.count_nonzero(lc2 != lc)
    assert changed <= (lc == nc).sum() * 4097
    # a 2-D red-black grid cannot lose a colour
    hexm = M.hex_mesh(4, 4, 4, seed=0, permute=False)
    nch, lch, _, _ = L and (lambda r: r)(__import__("ucfd_sparse_amd.layout",
                                                    fromlist=["x"]).greedy_coloring(hexm.nei_ele))
    nr, lr, _, _ = reduce_colors(hexm.nei_ele, lch)
    assert nch == 2 and nr == 2


def _bfs_dist(nei, seeds):
    """Independent breadth-first distances (and component roots) of the
    symmetrised face graph; distances do not depend on neighbour order."""
    from collections import deque
    F, N = nei.shape
    adj = [set() for _ in range(N)]
    for f in range(F):
        for i in range(N):
            j = int(nei[f, i])
            if j != i:
                adj[i].add(j)
                adj[j].add(i)
    dist = -np.ones(N, np.int64)
    root = -np.ones(N, np.int64)
    for s in seeds:
        if dist[s] >= 0:
            continue
        dist[s], root[s] = 0, s
        q = deque([s])
        while q:
            i = q.popleft()
            for j in adj[i]:
                if dist[j] < 0:
                    dist[j], root[j] = dist[i] + 1, s
                    q.append(j)
    return dist, root


@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_bfs_order_and_two_colouring(kind):
    from ucfd_sparse_amd.layout import bfs_order, colored_layout, greedy_coloring_ordered
    mesh = M.tet_mesh(6, seed=4) if kind == "tet" else M.hex_mesh(7, 6, 5, seed=4)
    nei = np.ascontiguousarray(mesh.nei_ele)
    F, N = nei.shape
    seeds = np.random.default_rng(5).permutation(N).astype(np.int32)
    order = bfs_order(nei, seeds)
    assert np.array_equal(np.sort(order), np.arange(N))
    dist, root = _bfs_dist(nei, seeds)
    # components are contiguous blocks entered at their root, distances
    # non-decreasing inside each block: a breadth-first order
    d, r = dist[order], root[order]
    starts = np.flatnonzero(np.r_[True, r[1:] != r[:-1]])
    assert np.array_equal(order[starts], r[starts]) and len(set(r[starts])) == len(starts)
    for a, b in zip(starts, list(starts[1:]) + [N]):
        assert np.all(np.diff(d[a:b]) >= 0)
    # the Kuhn split and the hex grid are bipartite: first-fit in this order
    # gives exactly two conflict-free colours
    nc, lc, ic, off = greedy_coloring_ordered(nei, order)
    assert nc == 2
    assert L.load().ucfd_coloring_conflicts(N, F, nei.ctypes.data, lc.ctypes.data) == 0
    lay = colored_layout(mesh, None, coloring="greedy-bfs")
    assert lay.ncolors == 2 and abs(int(np.diff(lay.off)[0]) - N // 2) <= N % 2 + 1
    bad = seeds.copy()
    bad[0] = bad[1]
    with pytest.raises(ValueError):
        bfs_order(nei, bad)


def test_bfs_greedy_on_an_odd_cycle_needs_three_colours():
    from ucfd_sparse_amd.layout import bfs_order, greedy_coloring_ordered
    # 5-cycle plus a boundary self-face column: not bipartite
    N = 5
    nei = np.array([[(i + 1) % N for i in range(N)], [(i - 1) % N for i in range(N)],
                    list(range(N))], np.int32)
    order = bfs_order(nei)
    assert order[0] == 0 and np.array_equal(np.sort(order), np.arange(N))
    nc, lc, _, _ = greedy_coloring_ordered(nei, order)
    assert nc == 3
    assert L.load().ucfd_coloring_conflicts(N, 3, nei.ctypes.data, lc.ctypes.data) == 0

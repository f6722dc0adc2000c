"""Host-side colouring and colour-contiguous renumbering (wrappers of the
library's host routines ``ucfd_greedy_coloring`` / ``ucfd_color_reorder`` /
``ucfd_permute_*``, include/ucfd_ext.h).

The colour-contiguous layout is how the MI355X path wants the mesh in HBM:
after renumbering, colour c occupies cells [off[c], off[c+1]), icolor is the
identity, every per-cell array is read coalesced by the sweep kernels and
the neighbour gathers stay local when the locality key (a Morton code of
the cell position) orders the cells inside each colour.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import lib as L
from .mesh import Flow, Mesh


def greedy_coloring(nei_ele: np.ndarray):
    """First-fit greedy colouring; returns (ncolors, lcolor, icolor, color_off)."""
    nei = np.ascontiguousarray(nei_ele, np.int32)
    F, N = nei.shape
    lcolor = np.empty(N, np.int32)
    icolor = np.empty(N, np.int32)
    off = np.empty(max(2 * F + 4, 64), np.int32)
    nc = L.load().ucfd_greedy_coloring(N, F, nei.ctypes.data, icolor.ctypes.data,
                                       lcolor.ctypes.data, off.ctypes.data)
    if nc <= 0 and N > 0:
        raise ValueError("colouring failed")
    return nc, lcolor, icolor, off[:nc + 1].copy()


def greedy_coloring_ordered(nei_ele: np.ndarray, order: np.ndarray):
    """First-fit greedy visiting cells in `order`; returns (ncolors, lcolor,
    icolor, color_off)."""
    nei = np.ascontiguousarray(nei_ele, np.int32)
    F, N = nei.shape
    o = np.ascontiguousarray(order, np.int32)
    lcolor = np.empty(N, np.int32)
    icolor = np.empty(N, np.int32)
    off = np.empty(max(2 * F + 4, 64), np.int32)
    nc = L.load().ucfd_greedy_coloring_ordered(N, F, nei.ctypes.data, o.ctypes.data,
                                               icolor.ctypes.data, lcolor.ctypes.data,
                                               off.ctypes.data)
    if nc <= 0 and N > 0:
        raise ValueError("colouring failed")
    return nc, lcolor, icolor, off[:nc + 1].copy()


def bfs_order(nei_ele: np.ndarray, seeds: np.ndarray | None = None) -> np.ndarray:
    """Breadth-first visiting order of the face graph (ucfd_bfs_order),
    components entered at their first cell in `seeds` (ascending id if None)."""
    nei = np.ascontiguousarray(nei_ele, np.int32)
    F, N = nei.shape
    order = np.empty(N, np.int32)
    sd = None if seeds is None else np.ascontiguousarray(seeds, np.int32)
    rc = L.load().ucfd_bfs_order(N, F, nei.ctypes.data, None if sd is None else sd.ctypes.data,
                                 order.ctypes.data)
    if rc < 0:
        raise ValueError("ucfd_bfs_order: bad input")
    return order


def reduce_colors(nei_ele: np.ndarray, lcolor: np.ndarray, max_chain: int = 4096):
    """Try to remove the highest colour by Kempe-chain swaps
    (ucfd_reduce_colors); returns (ncolors, lcolor, icolor, color_off)."""
    nei = np.ascontiguousarray(nei_ele, np.int32)
    F, N = nei.shape
    lc = np.array(lcolor, np.int32)
    icolor = np.empty(N, np.int32)
    off = np.empty(int(lc.max()) + 2 if N else 2, np.int32)
    nc = L.load().ucfd_reduce_colors(N, F, nei.ctypes.data, int(max_chain), icolor.ctypes.data,
                                     lc.ctypes.data, off.ctypes.data)
    if nc <= 0 and N > 0:
        raise ValueError("colour reduction failed")
    return nc, lc, icolor, off[:nc + 1].copy()


def balanced_coloring(nei_ele: np.ndarray, ncolors: int = 0):
    """Balanced colouring (ucfd_balanced_coloring); returns (ncolors, lcolor,
    icolor, color_off)."""
    nei = np.ascontiguousarray(nei_ele, np.int32)
    F, N = nei.shape
    lcolor = np.empty(N, np.int32)
    icolor = np.empty(N, np.int32)
    off = np.empty(max(2 * F + 4, ncolors + 2, 64), np.int32)
    nc = L.load().ucfd_balanced_coloring(N, F, nei.ctypes.data, int(ncolors),
                                         icolor.ctypes.data, lcolor.ctypes.data, off.ctypes.data)
    if nc <= 0 and N > 0:
        raise ValueError("colouring failed")
    return nc, lcolor, icolor, off[:nc + 1].copy()


def color_reorder(lcolor: np.ndarray, key: np.ndarray | None = None):
    """perm[old] = new, sorted by (colour, key, id); returns (perm, color_off)."""
    lc = np.ascontiguousarray(lcolor, np.int32)
    N = lc.shape[0]
    nc = int(lc.max()) if N else 0
    perm = np.empty(N, np.int32)
    off = np.empty(nc + 1, np.int32)
    kp = None
    if key is not None:
        key = np.ascontiguousarray(key, np.uint64)
        kp = key.ctypes.data
    L.load().ucfd_color_reorder(N, lc.ctypes.data, kp, perm.ctypes.data, off.ctypes.data)
    return perm, off


def permute_cells(a: np.ndarray, perm: np.ndarray) -> np.ndarray:
    """out[..., perm[i]] = a[..., i]"""
    out = np.empty_like(a)
    out[..., perm] = a
    return out


def permute_nei(nei_ele: np.ndarray, perm: np.ndarray) -> np.ndarray:
    nei = np.ascontiguousarray(nei_ele, np.int32)
    F, N = nei.shape
    out = np.empty_like(nei)
    p = np.ascontiguousarray(perm, np.int32)
    L.load().ucfd_permute_nei(N, F, p.ctypes.data, nei.ctypes.data, out.ctypes.data)
    return out


@dataclass
class ColoredLayout:
    mesh: Mesh            # renumbered: colour c = cells off[c]..off[c+1]-1
    flow: Flow | None
    lcolor: np.ndarray    # non-decreasing
    icolor: np.ndarray    # identity
    off: np.ndarray       # [ncolors+1]
    perm: np.ndarray      # perm[original] = new

    @property
    def ncolors(self) -> int:
        return len(self.off) - 1


def colored_layout(mesh: Mesh, flow: Flow | None = None, use_key: bool = True,
                   coloring: str = "greedy", ncolors: int = 0) -> ColoredLayout:
    """Colour the mesh and renumber it colour-contiguously (locality-key order
    inside each colour when use_key).  coloring: "greedy" (first-fit in
    ascending cell id), "greedy-key" (first-fit visiting cells in locality-key
    order), "greedy-key-reduced" (the same, then the highest colour removed
    by Kempe-chain swaps where possible), "greedy-bfs" (first-fit in
    breadth-first order from the lowest locality key: the 2-colouring of a
    bipartite face graph, then Kempe-chain reduction) or "balanced"."""
    if coloring == "balanced":
        nc, lcolor, icolor, off = balanced_coloring(mesh.nei_ele, ncolors)
    elif coloring == "greedy-bfs":
        seeds = np.argsort(np.asarray(mesh.key, np.uint64), kind="stable").astype(np.int32)
        order = bfs_order(mesh.nei_ele, seeds)
        nc, lcolor, icolor, off = greedy_coloring_ordered(mesh.nei_ele, order)
        if nc > 2:
            nc, lcolor, icolor, off = reduce_colors(mesh.nei_ele, lcolor)
    elif coloring in ("greedy-key", "greedy-key-reduced"):
        order = np.argsort(np.asarray(mesh.key, np.uint64), kind="stable").astype(np.int32)
        nc, lcolor, icolor, off = greedy_coloring_ordered(mesh.nei_ele, order)
        if coloring == "greedy-key-reduced":
            nc, lcolor, icolor, off = reduce_colors(mesh.nei_ele, lcolor)
    else:
        nc, lcolor, icolor, off = greedy_coloring(mesh.nei_ele)
    perm, off2 = color_reorder(lcolor, mesh.key if use_key else None)
    assert np.array_equal(off, off2)
    p = lambda a: None if a is None else permute_cells(a, perm)  # noqa: E731
    m2 = Mesh(permute_nei(mesh.nei_ele, perm), p(mesh.fnorm_vol), p(mesh.vec_fnorm),
              p(mesh.orig_id), p(mesh.key), vol=p(mesh.vol), centroid=p(mesh.centroid),
              face_area=p(mesh.face_area))
    f2 = None if flow is None else Flow(p(flow.upts), p(flow.fspr), p(flow.dt), p(flow.rhsb))
    return ColoredLayout(m2, f2, p(lcolor), np.arange(mesh.neles, dtype=np.int32), off, perm)

"""Seeded synthetic meshes and flow fields (host side, numpy arrays).

Thin ctypes wrapper over ``libucfdmesh.so`` (``csrc/meshgen.c``), the one C
generator shared by the tests, the golden-vector script and ``bench.py``, so
that every consumer sees bit-identical inputs for a given seed.

Layouts follow the legacy LU-SGS array API (``examples/omp3d.c:111-121``):
``nei_ele[nface, N]`` int32, ``fnorm_vol[nface, N]``, ``fspr[nface, N]``,
``vec_fnorm[nface, ndims, N]``, ``upts/rhsb[nvars, N]``, ``dt[N]``.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libucfdmesh.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build() "
                               "or make -C ucfd_sparse_amd/csrc")
        lib = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        lib.ucfdgen_hex.argtypes = [ctypes.c_int] * 3 + [ctypes.c_uint64, ctypes.c_int,
                                                          ctypes.c_double, vp, vp, vp, vp, vp]
        lib.ucfdgen_hex.restype = ctypes.c_int
        lib.ucfdgen_kuhn_tets.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int,
                                          ctypes.c_double] + [vp] * 8
        lib.ucfdgen_kuhn_tets.restype = ctypes.c_int
        lib.ucfdgen_flow.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_uint64, vp, vp, vp, vp, vp, vp]
        lib.ucfdgen_flow.restype = ctypes.c_int
        lib.ucfdgen_permutation.argtypes = [ctypes.c_int64, ctypes.c_uint64, ctypes.c_int, vp]
        lib.ucfdgen_permutation.restype = None
        lib.ucfdgen_face_pairs.argtypes = [ctypes.c_int64, ctypes.c_int, vp, vp]
        lib.ucfdgen_face_pairs.restype = ctypes.c_int64
        lib.ucfdgen_hilbert_keys.argtypes = [ctypes.c_int] * 4 + [ctypes.c_int64, vp, vp]
        lib.ucfdgen_hilbert_keys.restype = None
        lib.ucfdgen_normal_fill.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64, vp]
        lib.ucfdgen_normal_fill.restype = None
        lib.ucfdgen_bsr_count.argtypes = [ctypes.c_int64, ctypes.c_int, vp]
        lib.ucfdgen_bsr_count.restype = ctypes.c_int64
        lib.ucfdgen_bsr_fill.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_uint64] + [vp] * 10
        lib.ucfdgen_bsr_fill.restype = ctypes.c_int
        lib.ucfdgen_flowsys_fill.argtypes = ([ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_uint64] + [vp] * 16
                                             + [ctypes.c_int])
        lib.ucfdgen_flowsys_fill.restype = ctypes.c_int
        lib.ucfdgen_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
        lib.ucfdgen_uniform.restype = ctypes.c_double
        lib.ucfdgen_normal.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
        lib.ucfdgen_normal.restype = ctypes.c_double
        _LIB = lib
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data


@dataclass
class Mesh:
    """An unstructured cell mesh in the legacy nface-major layout."""
    nei_ele: np.ndarray          # int32 [nface, N]
    fnorm_vol: np.ndarray        # f64 [nface, N]  face area / cell volume
    vec_fnorm: np.ndarray        # f64 [nface, ndims, N] outward unit normals
    orig_id: np.ndarray          # int32 [N] id before the seeded permutation
    key: np.ndarray              # uint64 [N] Morton key of the cell position
    vol: np.ndarray | None = None
    centroid: np.ndarray | None = None     # f64 [ndims, N]
    face_area: np.ndarray | None = None    # f64 [nface, N]
    extra: dict = field(default_factory=dict)

    @property
    def neles(self) -> int:
        return self.nei_ele.shape[1]

    @property
    def nface(self) -> int:
        return self.nei_ele.shape[0]

    @property
    def ndims(self) -> int:
        return self.vec_fnorm.shape[1]


def hex_mesh(m: int, n: int, l: int, seed: int = 0, permute: bool = True,
             voljit: float = 0.1) -> Mesh:
    """Unit-cube m x n x l hex grid, faces -x,+x,-y,+y,-z,+z, boundary -> self
    (``examples/utils/pbutils.c:43-84``), with a seeded cell permutation and
    the cell volume jittered by U[1-voljit, 1+voljit]."""
    N = m * n * l
    nei = np.empty((6, N), np.int32)
    fnv = np.empty((6, N), np.float64)
    vfn = np.empty((6, 3, N), np.float64)
    oid = np.empty(N, np.int32)
    key = np.empty(N, np.uint64)
    rc = _lib().ucfdgen_hex(m, n, l, seed, int(permute), voljit, _p(nei), _p(fnv), _p(vfn),
                            _p(oid), _p(key))
    if rc:
        raise ValueError(f"ucfdgen_hex failed ({rc})")
    return Mesh(nei, fnv, vfn, oid, key)


def tet_mesh(m: int, seed: int = 0, permute: bool = True, jitter: float = 0.1) -> Mesh:
    """m^3 hex grid split into 6 Kuhn tets per hex (N = 6 m^3, nface = 4),
    vertices jittered +-jitter*h (tangential only on the boundary), normals,
    areas and volumes from the coordinates, seeded cell permutation."""
    N = 6 * m ** 3
    nei = np.empty((4, N), np.int32)
    fnv = np.empty((4, N), np.float64)
    vfn = np.empty((4, 3, N), np.float64)
    vol = np.empty(N, np.float64)
    cen = np.empty((3, N), np.float64)
    oid = np.empty(N, np.int32)
    key = np.empty(N, np.uint64)
    area = np.empty((4, N), np.float64)
    rc = _lib().ucfdgen_kuhn_tets(m, seed, int(permute), jitter, _p(nei), _p(fnv), _p(vfn),
                                  _p(vol), _p(cen), _p(oid), _p(key), _p(area))
    if rc:
        raise ValueError(f"ucfdgen_kuhn_tets failed ({rc})")
    return Mesh(nei, fnv, vfn, oid, key, vol=vol, centroid=cen, face_area=area)


@dataclass
class Flow:
    upts: np.ndarray    # [nvars, N]
    fspr: np.ndarray    # [nface, N]
    dt: np.ndarray      # [N]
    rhsb: np.ndarray    # [nvars, N]


def flow_fields(mesh: Mesh, seed: int = 0, nvars: int | None = None) -> Flow:
    """Euler states and LU-SGS inputs: rho~U[0.8,1.2], u,v,w~U[-0.3,0.3],
    p~U[0.7,1.3], fspr = |u.n| + c, dt~U[1e-3,1e-2], rhsb~N(0,1e-3)."""
    N, F, D = mesh.neles, mesh.nface, mesh.ndims
    nv = nvars if nvars is not None else D + 2
    upts = np.empty((nv, N), np.float64)
    rhsb = np.empty((nv, N), np.float64)
    fspr = np.empty((F, N), np.float64)
    dt = np.empty(N, np.float64)
    rc = _lib().ucfdgen_flow(N, F, D, nv, seed, _p(mesh.orig_id), _p(mesh.vec_fnorm),
                             _p(upts), _p(fspr), _p(dt), _p(rhsb))
    if rc:
        raise ValueError(f"ucfdgen_flow failed ({rc})")
    return Flow(upts, fspr, dt, rhsb)


def hilbert_keys(mesh: Mesh, kind: str, dims, tet_type_major: bool = False) -> np.ndarray:
    """Hilbert-curve locality key of every cell (from its original id):
    kind "hex" with dims (m, n, l), or "tet" with dims (m,) for Kuhn tets
    (key = 8*hilbert(cube) + tet index).  tet_type_major: tets ordered by
    their index in the cube first, then along the curve (key = tet index *
    2^40 + hilbert(cube)); in a colour-contiguous layout the 64 lanes of a
    wave then cover 64 cubes, and a face's neighbours of consecutive cells
    fall in fewer cache lines."""
    key = np.empty(mesh.neles, np.uint64)
    d = tuple(dims) + (dims[0],) * (3 - len(dims))
    _lib().ucfdgen_hilbert_keys(0 if kind == "hex" else 1, d[0], d[1], d[2], mesh.neles,
                                _p(np.ascontiguousarray(mesh.orig_id, np.int32)), _p(key))
    if kind == "tet" and tet_type_major:
        key = ((key % np.uint64(8)) << np.uint64(40)) | (key // np.uint64(8))
    return key


def face_pairs(nei_ele: np.ndarray) -> np.ndarray:
    """pair[f, i] = face f' of nb = nei_ele[f, i] with nei_ele[f', nb] == i
    (-1 on boundary faces)."""
    nei = np.ascontiguousarray(nei_ele, np.int32)
    F, N = nei.shape
    pair = np.empty((F, N), np.int32)
    bad = _lib().ucfdgen_face_pairs(N, F, _p(nei), _p(pair))
    if bad:
        raise ValueError(f"{bad} interior faces have no back face")
    return pair


def permutation(n: int, seed: int, permute: bool = True) -> np.ndarray:
    perm = np.empty(n, np.int32)
    _lib().ucfdgen_permutation(n, seed, int(permute), _p(perm))
    return perm


def uniform(seed: int, stream: int, idx: np.ndarray) -> np.ndarray:
    f = _lib().ucfdgen_uniform
    return np.array([f(seed, stream, int(i)) for i in np.ravel(idx)])


def normal(seed: int, stream: int, idx: np.ndarray) -> np.ndarray:
    f = _lib().ucfdgen_normal
    return np.array([f(seed, stream, int(i)) for i in np.ravel(idx)])


@dataclass
class BSR:
    """Block sparse row matrix, row-major blocks (the layout of
    src/krylov/matrix/sparsemat.c:57-101)."""
    bn: int
    blk: int
    rowptr: np.ndarray       # int32 [bn+1]
    colidx: np.ndarray       # int32 [nnz]
    diagslots: np.ndarray    # int32 [bn] index of the diagonal block of each row
    values: np.ndarray       # f64 [nnz, blk, blk]

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.bsr_matrix((self.values, self.colidx, self.rowptr),
                             shape=(self.bn * self.blk, self.bn * self.blk))


def block_system(mesh: Mesh, flow: Flow, blk: int = 5, seed: int = 0) -> BSR:
    """The block Jacobian of the block LU-SGS and Krylov workloads on a mesh
    (definition in ucfdgen_bsr_fill, csrc/meshgen.c)."""
    N, F = mesh.neles, mesh.nface
    lib = _lib()
    nnz = lib.ucfdgen_bsr_count(N, F, _p(mesh.nei_ele))
    rowptr = np.empty(N + 1, np.int32)
    colidx = np.empty(nnz, np.int32)
    diag = np.empty(N, np.int32)
    values = np.empty((nnz, blk, blk), np.float64)
    rc = lib.ucfdgen_bsr_fill(N, F, blk, seed, _p(mesh.nei_ele), _p(mesh.fnorm_vol),
                              _p(flow.fspr), _p(flow.dt), _p(mesh.orig_id),
                              _p(mesh.centroid), _p(rowptr), _p(colidx), _p(diag), _p(values))
    if rc:
        raise ValueError("ucfdgen_bsr_fill failed")
    return BSR(N, blk, rowptr, colidx, diag, values)


def normal_fill(n: int, seed: int, stream: int = 77) -> np.ndarray:
    out = np.empty(n, np.float64)
    _lib().ucfdgen_normal_fill(seed, stream, n, _p(out))
    return out


@dataclass
class FlowSysMesh:
    """Face-slot CSR arrays of the FlowSys API (src/lusgs/flowsys.h:34-59)."""
    nlocal: int
    nfaces: int              # number of face slots
    rowptr: np.ndarray       # int32 [nlocal+1]
    colidx: np.ndarray       # int32 [npos] neighbour cell of each position
    slots: np.ndarray        # int32 [npos] face slot of each position
    sides: np.ndarray        # int8 [npos] +1 owner side, -1 other side
    face_area: np.ndarray    # f64 [nfaces]
    face_normal: np.ndarray  # f64 [ndims, nfaces]
    rcp_vol: np.ndarray      # f64 [nlocal]
    jmat: np.ndarray         # f64 [2, nv, nv, nfaces]


def flowsys_system(mesh: Mesh, flow: Flow, nv: int = 5, seed: int = 0) -> FlowSysMesh:
    """FlowSys arrays of a mesh (definition in ucfdgen_flowsys_fill,
    csrc/meshgen.c): interior faces as positions, jmat = A+/A- of the
    block Jacobian used by block_system.  Meshes without volumes (the hex
    generator) use rcp_vol = 1 and the owner's fnorm_vol as face area."""
    N, F = mesh.neles, mesh.nface
    D = mesh.vec_fnorm.shape[1]
    lib = _lib()
    ns, npos = ctypes.c_int64(), ctypes.c_int64()
    area = mesh.face_area if mesh.vol is not None else mesh.fnorm_vol
    rcp = 1.0 / mesh.vol if mesh.vol is not None else np.ones(N)
    args = [N, F, D, nv, seed, _p(mesh.nei_ele), _p(area), _p(mesh.vec_fnorm),
            _p(flow.fspr), _p(mesh.orig_id), _p(mesh.centroid), _p(mesh.fnorm_vol)]
    rc = lib.ucfdgen_flowsys_fill(*args, ctypes.byref(ns), ctypes.byref(npos), *([None] * 7), 0)
    if rc:
        raise ValueError("ucfdgen_flowsys_fill failed")
    ns, npos = ns.value, npos.value
    rowptr = np.empty(N + 1, np.int32)
    colidx = np.empty(npos, np.int32)
    slots = np.empty(npos, np.int32)
    sides = np.empty(npos, np.int8)
    face_area = np.empty(ns)
    face_normal = np.empty((D, ns))
    jmat = np.empty((2, nv, nv, ns))
    rc = lib.ucfdgen_flowsys_fill(*args, ctypes.byref(ctypes.c_int64()),
                                  ctypes.byref(ctypes.c_int64()), _p(rowptr), _p(colidx),
                                  _p(slots), _p(sides), _p(face_area), _p(face_normal), _p(jmat), 1)
    if rc:
        raise ValueError("ucfdgen_flowsys_fill failed")
    return FlowSysMesh(N, ns, rowptr, colidx, slots, sides, face_area, face_normal, rcp, jmat)


def flowsys_elements(mesh: Mesh, flow: Flow, nvars: int = 6, seed: int = 0) -> dict:
    """One element type covering every cell of the mesh for the FlowSys API:
    cell_ids = a seeded permutation (element-local index -> rank cell), and
    element arrays in [nvars, neles] layout: uptsb (the flow state, then
    turbulence variables U[0.1, 1]), rhs ~ N(0, 1e-3), dt (the flow's),
    dsrc ~ U[0, 1] (source-Jacobian inputs), vol (the cells' volumes)."""
    N = mesh.neles
    rng = np.random.default_rng(seed + 4242)
    cell_ids = rng.permutation(N).astype(np.int32)
    nf = flow.upts.shape[0]
    up = np.empty((nvars, N))
    up[:min(nf, nvars)] = flow.upts[:min(nf, nvars), cell_ids]
    if nvars > nf:
        up[nf:] = rng.uniform(0.1, 1.0, size=(nvars - nf, N))
    return {"cell_ids": cell_ids, "uptsb": up,
            "rhs": rng.normal(0.0, 1e-3, size=(nvars, N)),
            "dt": np.ascontiguousarray(flow.dt[cell_ids]),
            "dsrc": rng.uniform(0.0, 1.0, size=(nvars, N)),
            "vol": (np.ascontiguousarray(mesh.vol[cell_ids]) if mesh.vol is not None
                    else np.ones(N))}

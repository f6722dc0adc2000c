"""Geometric invariants of the seeded mesh generators (csrc/meshgen.c): the
Kuhn-tet meshes of configs 1-4 take their normals, areas and volumes from
jittered vertex coordinates, so they must be closed and conforming; the hex
grid of config 0 carries exact axis normals in the neighbour table's face
order (examples/utils/pbutils.c:27-32; boundary faces point to the cell
itself).  A wrong sign, a transposed face index or a missing back face in
the generator breaks one of these."""
import numpy as np
import pytest

from ucfd_sparse_amd import mesh as M


@pytest.fixture(scope="module", params=[(4, 7, True), (6, 11, True), (5, 3, False)])
def tet(request):
    m, seed, permute = request.param
    return m, M.tet_mesh(m, seed=seed, permute=permute)


def back_faces(nei):
    """pair[f, i] = f' with nei[f', nb] == i for interior faces, -1 on the
    boundary; raises if a back face is missing."""
    F, N = nei.shape
    pair = np.full((F, N), -1, np.int64)
    for f in range(F):
        nb = nei[f]
        inner = nb != np.arange(N)
        hits = np.stack([nei[g, nb] == np.arange(N) for g in range(F)])   # [F', N]
        ok = hits.any(axis=0)
        assert np.all(ok[inner]), f"face {f}: {np.sum(~ok & inner)} faces without a back face"
        pair[f, inner] = np.argmax(hits, axis=0)[inner]
    return pair


def test_tet_cells_are_closed(tet):
    """sum_f area_f n_f = 0 for every cell (the divergence theorem for a
    constant field)."""
    m, mesh = tet
    area = mesh.face_area
    closure = np.einsum("fn,fdn->dn", area, mesh.vec_fnorm)
    scale = area.max()
    assert np.abs(closure).max() <= 1e-12 * scale
    # the stored ratio is area / volume
    assert np.allclose(mesh.fnorm_vol, area / mesh.vol, rtol=1e-14, atol=0)


def test_tet_volumes_fill_the_unit_cube(tet):
    m, mesh = tet
    assert np.all(mesh.vol > 0)
    assert abs(mesh.vol.sum() - 1.0) <= 1e-12
    assert mesh.neles == 6 * m ** 3


def test_tet_normals_are_unit_and_outward(tet):
    m, mesh = tet
    n = mesh.vec_fnorm
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, rtol=0, atol=1e-13)
    nei, c = mesh.nei_ele, mesh.centroid
    F, N = nei.shape
    ids = np.arange(N)
    for f in range(F):
        nb = nei[f]
        inner = nb != ids
        # the neighbour's centroid lies on the outer side of the face
        d = c[:, nb] - c
        assert np.all(np.einsum("dn,dn->n", d[:, inner], n[f][:, inner]) > 0)
        # boundary faces lie on the unit cube's walls and point out of it
        b = ~inner
        axis = np.argmax(np.abs(n[f][:, b]), axis=0)
        comp = n[f][axis, np.nonzero(b)[0]]
        assert np.allclose(np.abs(comp), 1.0, atol=1e-13)
        cen = c[axis, np.nonzero(b)[0]]
        assert np.all((comp > 0) == (cen > 0.5))


def test_tet_neighbours_are_symmetric_with_opposite_normals(tet):
    m, mesh = tet
    nei = mesh.nei_ele
    F, N = nei.shape
    pair = back_faces(nei)
    ids = np.arange(N)
    for f in range(F):
        nb = nei[f]
        inner = nb != ids
        fp = pair[f, inner]
        j = nb[inner]
        assert np.array_equal(nei[fp, j], ids[inner])
        assert np.allclose(mesh.vec_fnorm[fp, :, j], -mesh.vec_fnorm[f][:, inner].T,
                           rtol=0, atol=1e-13)
        assert np.allclose(mesh.face_area[fp, j], mesh.face_area[f, inner], rtol=1e-13, atol=0)


def test_tet_generator_matches_its_face_pair_map(tet):
    m, mesh = tet
    assert np.array_equal(M.face_pairs(mesh.nei_ele), back_faces(mesh.nei_ele))


@pytest.mark.parametrize("dims, permute", [((5, 4, 3), True), ((4, 4, 2), False)])
def test_hex_grid_axis_normals_and_symmetry(dims, permute):
    mesh = M.hex_mesh(*dims, seed=9, permute=permute)
    nei, n = mesh.nei_ele, mesh.vec_fnorm
    F, N = nei.shape
    assert F == 6 and N == dims[0] * dims[1] * dims[2]
    axes = np.array([[-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0], [0, 0, -1], [0, 0, 1]], float)
    for f in range(6):
        assert np.array_equal(n[f], np.repeat(axes[f][:, None], N, axis=1))
    pair = back_faces(nei)
    ids = np.arange(N)
    for f in range(6):
        inner = nei[f] != ids
        # the back face of -x is +x and so on
        assert np.all(pair[f, inner] == (f ^ 1))
    # boundary faces: as many as the grid's outer faces
    nb_faces = sum(int(np.sum(nei[f] == ids)) for f in range(6))
    m, k, l = dims
    assert nb_faces == 2 * (m * k + k * l + m * l)
    # fnorm_vol = area / (volume x U[0.9, 1.1]) = (1/h_axis) / U[0.9, 1.1],
    # h = 1/m, 1/n, 1/l along x, y, z; both faces of an axis share the draw
    h = np.repeat(1.0 / np.array(dims, float), 2)
    r = mesh.fnorm_vol * h[:, None]
    assert np.all((r > 1 / 1.1 - 1e-12) & (r < 1 / 0.9 + 1e-12))
    assert np.array_equal(r[0], r[1]) and np.allclose(r[0], r[2], rtol=1e-14, atol=0)

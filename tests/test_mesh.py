"""Mesh builder: the reference's mesh fixtures, brute-force connectivity and
geometric invariants (CPU only)."""

import numpy as np
import pytest

from anuga_core_amd.mesh import Mesh, build_neighbour_structure, rectangular, rectangular_cross

# 6-point / 4-triangle fixture of anuga/abstract_2d_finite_volumes/tests/test_neighbour_mesh.py:317
A, B, C, D, E, F = [0.0, 0.0], [0.0, 2.0], [2.0, 0.0], [0.0, 4.0], [2.0, 2.0], [4.0, 0.0]
POINTS = [A, B, C, D, E, F]
VERTICES = [[1, 0, 2], [1, 2, 4], [4, 2, 5], [3, 1, 4]]   # bac, bce, ecf, dbe


def test_reference_neighbour_fixture():
    """test_mesh_and_neighbours (test_neighbour_mesh.py:317-380)."""
    m = Mesh(POINTS, VERTICES)
    assert list(m.number_of_boundaries) == [2, 0, 2, 2]
    nb, ne = m.neighbours, m.neighbour_edges
    assert nb[0, 0] < 0 and nb[0, 1] == 1 and nb[0, 2] < 0
    assert list(nb[1]) == [2, 3, 0]
    assert nb[2, 0] < 0 and nb[2, 1] < 0 and nb[2, 2] == 1
    assert nb[3, 0] == 1 and nb[3, 1] < 0 and nb[3, 2] < 0
    assert ne[0, 1] == 2 and list(ne[1]) == [2, 0, 1] and ne[2, 2] == 0 and ne[3, 0] == 1


def test_reference_surrogate_fixture():
    """test_surrogate_neighbours (test_neighbour_mesh.py:489-530)."""
    m = Mesh(POINTS, VERTICES)
    s = m.surrogate_neighbours
    assert list(s[0]) == [0, 1, 0]
    assert list(s[1]) == [2, 3, 0]
    assert list(s[2]) == [2, 2, 1]
    assert list(s[3]) == [1, 3, 3]


def test_reference_areas_and_midpoints():
    """test_areas (test_general_mesh.py:348) and edge midpoints (:127-150)."""
    p, v, _ = rectangular(1, 3)
    m = Mesh(p, v)
    assert m.areas.sum() == 1.0
    E = m.edge_midpoint_coordinates
    for i in range(m.number_of_triangles):
        k0, k1, k2 = v[i]
        assert np.allclose(E[3 * i + 0], 0.5 * (p[k1] + p[k2]))
        assert np.allclose(E[3 * i + 1], 0.5 * (p[k0] + p[k2]))
        assert np.allclose(E[3 * i + 2], 0.5 * (p[k1] + p[k0]))


def test_degenerate_triangle_rejected():
    pts = np.array([[1.0, 1.0], B, C, D, E, F])
    with pytest.raises(ValueError):
        Mesh(pts, VERTICES)


def _brute_neighbours(tri):
    """O(N^2) edge matching: edge i of k is (v[i+1], v[i+2]); the neighbour holds
    the same two vertices."""
    N = len(tri)
    nb = -np.ones((N, 3), np.int64)
    ne = -np.ones((N, 3), np.int64)
    for k in range(N):
        for i in range(3):
            a, b = tri[k][(i + 1) % 3], tri[k][(i + 2) % 3]
            for n in range(N):
                if n == k:
                    continue
                for j in range(3):
                    c, d = tri[n][(j + 1) % 3], tri[n][(j + 2) % 3]
                    if (a, b) == (d, c):
                        nb[k, i], ne[k, i] = n, j
    return nb, ne


@pytest.mark.parametrize("factory,m,n", [(rectangular_cross, 3, 2), (rectangular, 4, 3),
                                         (rectangular_cross, 1, 1)])
def test_neighbours_brute_force(factory, m, n):
    p, v, b = factory(m, n)
    nb, ne, nob = build_neighbour_structure(len(p), v)
    bnb, bne = _brute_neighbours(v.tolist())
    assert np.array_equal(nb, bnb)
    assert np.array_equal(ne, bne)
    assert np.array_equal(nob, (bnb < 0).sum(axis=1))


def test_rectangular_cross_layout():
    """Point/triangle order and tags of mesh_factory_ext.pyx:9-90."""
    m, n, L1, L2 = 3, 2, 3.0, 4.0
    p, v, b = rectangular_cross(m, n, L1, L2)
    assert p.shape == ((m + 1) * (n + 1) + m * n, 2)
    assert v.shape == (4 * m * n, 3)
    # grid point (i, j) at index i*(n+1)+j
    assert np.array_equal(p[1 * (n + 1) + 2], [1.0, 4.0])
    # first cell (0,0): left, bottom, right, top triangles around the centre
    v1, v2, v3, v4, v5 = 1, 0, n + 2, n + 1, (m + 1) * (n + 1)
    assert v[:4].tolist() == [[v2, v5, v1], [v4, v5, v2], [v3, v5, v4], [v1, v5, v3]]
    assert np.allclose(p[v5], [0.5, 1.0])
    tags = {}
    for (vol, edge), t in b.items():
        assert edge == 1
        tags.setdefault(t, 0)
        tags[t] += 1
    assert tags == {"left": n, "right": n, "bottom": m, "top": m}
    mesh = Mesh(p, v, b)
    assert mesh.boundary_length == 2 * (m + n)
    assert abs(mesh.areas.sum() - L1 * L2) < 1e-12


def test_geometry_invariants():
    p, v, b = rectangular_cross(7, 5, 70.0, 30.0)
    m = Mesh(p, v, b)
    N = m.number_of_triangles
    nrm = m.normals.reshape(N, 3, 2)
    assert np.allclose(np.hypot(nrm[..., 0], nrm[..., 1]), 1.0)
    # outward: normal . (midpoint - centroid) > 0
    mid = m.edge_midpoint_coordinates.reshape(N, 3, 2)
    assert np.all(np.sum(nrm * (mid - m.centroid_coordinates[:, None, :]), axis=2) > 0)
    # across an interior edge the normal is exactly negated and the length equal
    nb = m.neighbours
    ne = m.neighbour_edges
    k, i = np.nonzero(nb >= 0)
    assert np.array_equal(nrm[k, i], -nrm[nb[k, i], ne[k, i]])
    assert np.array_equal(m.edgelengths[k, i], m.edgelengths[nb[k, i], ne[k, i]])
    # radius is the distance from the centroid to the nearest edge midpoint
    d = np.sqrt(((mid - m.centroid_coordinates[:, None, :]) ** 2).sum(axis=2))
    assert np.allclose(m.radii, d.min(axis=1))


def test_boundary_enumeration_sorted():
    """neighbour_mesh.py:441-505: boundary ids follow sorted (vol, edge)."""
    p, v, b = rectangular_cross(4, 3)
    m = Mesh(p, v, b)
    keys = sorted(b.keys())
    for j, (vol, edge) in enumerate(keys):
        assert m.neighbours[vol, edge] == -(j + 1)
        assert m.boundary_cells[j] == vol and m.boundary_edges[j] == edge
    for tag, ids in m.tag_boundary_cells.items():
        for j in ids:
            assert b[(m.boundary_cells[j], m.boundary_edges[j])] == tag


def test_default_boundary_tag():
    m = Mesh(POINTS, VERTICES)
    assert set(m.boundary.values()) == {"exterior"}
    assert m.boundary_length == 6

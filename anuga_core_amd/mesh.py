"""Triangular mesh generation and the static mesh arrays of the DE step.

Everything here is set-up work done once on the host with vectorised numpy, so a
16M-triangle `rectangular_cross` mesh builds in seconds. The arrays follow the
reference's data model exactly, including the floating-point operation order of
every geometric quantity, so that the step computed on them is bit-comparable:

* mesh factories: `rectangular_cross` (anuga/abstract_2d_finite_volumes/
  mesh_factory_ext.pyx:9-90, point/triangle order and boundary tags) and
  `rectangular` (mesh_factory.py:84-135);
* geometry: areas, normals, edge lengths, centroids, radii
  (general_mesh.py:160-285) and edge midpoints (general_mesh.py:599-628);
* connectivity: neighbours / neighbour_edges / number_of_boundaries with edge i
  opposite vertex i (neighbour_table.cpp:63-212), surrogate neighbours
  (neighbour_mesh.py:318-346), boundary dictionary completion with the default
  tag (neighbour_mesh_ext.pyx:9-31) and boundary enumeration by sorted
  (vol, edge) -> neighbours = -(j+1) (neighbour_mesh.py:441-505).
"""

from __future__ import annotations

import numpy as np

DEFAULT_BOUNDARY_TAG = "exterior"  # anuga/config.py:32


# ---------------------------------------------------------------------------
# Mesh factories
# ---------------------------------------------------------------------------

def _boundary_dict(vols, edges, tag, out):
    for v, e in zip(vols.tolist(), edges.tolist()):
        out[(v, e)] = tag


def rectangular_cross(m, n, len1=1.0, len2=1.0, origin=(0.0, 0.0)):
    """m x n rectangular cells, each split into 4 triangles around its centre.

    Returns (points, elements, boundary) like
    anuga/abstract_2d_finite_volumes/mesh_factory.py:138-175. Grid points come
    first in (i, j) row-major order, then one centre point per cell in cell
    order; each cell yields the triangles left, bottom, right, top
    (mesh_factory_ext.pyx:37-90).
    """
    m = int(m)
    n = int(n)
    len1 = float(len1)
    len2 = float(len2)
    ox, oy = float(origin[0]), float(origin[1])
    delta1 = len1 / m
    delta2 = len2 / n

    ng = (m + 1) * (n + 1)
    points = np.empty((ng + m * n, 2), dtype=np.float64)
    ii = np.repeat(np.arange(m + 1, dtype=np.float64), n + 1)
    jj = np.tile(np.arange(n + 1, dtype=np.float64), m + 1)
    points[:ng, 0] = ii * delta1 + ox
    points[:ng, 1] = jj * delta2 + oy

    ci = np.repeat(np.arange(m, dtype=np.int64), n)
    cj = np.tile(np.arange(n, dtype=np.int64), m)
    v1 = ci * (n + 1) + (cj + 1)
    v2 = ci * (n + 1) + cj
    v3 = (ci + 1) * (n + 1) + (cj + 1)
    v4 = (ci + 1) * (n + 1) + cj
    # centre = (p1 + p2 + p3 + p4) * 0.25, summed left to right
    for d in range(2):
        points[ng:, d] = (((points[v1, d] + points[v2, d]) + points[v3, d])
                          + points[v4, d]) * 0.25
    v5 = ng + np.arange(m * n, dtype=np.int64)

    elements = np.empty((4 * m * n, 3), dtype=np.int64)
    e = elements.reshape(m * n, 4, 3)
    e[:, 0, 0], e[:, 0, 1], e[:, 0, 2] = v2, v5, v1   # left
    e[:, 1, 0], e[:, 1, 1], e[:, 1, 2] = v4, v5, v2   # bottom
    e[:, 2, 0], e[:, 2, 1], e[:, 2, 2] = v3, v5, v4   # right
    e[:, 3, 0], e[:, 3, 1], e[:, 3, 2] = v1, v5, v3   # top

    cell = np.arange(m * n, dtype=np.int64)
    boundary = {}
    sel = ci == 0
    _boundary_dict(4 * cell[sel] + 0, np.ones(sel.sum(), np.int64), "left", boundary)
    sel = cj == 0
    _boundary_dict(4 * cell[sel] + 1, np.ones(sel.sum(), np.int64), "bottom", boundary)
    sel = ci == m - 1
    _boundary_dict(4 * cell[sel] + 2, np.ones(sel.sum(), np.int64), "right", boundary)
    sel = cj == n - 1
    _boundary_dict(4 * cell[sel] + 3, np.ones(sel.sum(), np.int64), "top", boundary)
    # the reference inserts tags cell by cell; key order is irrelevant because
    # boundary enumeration sorts the keys
    return points, elements, boundary


def rectangular(m, n, len1=1.0, len2=1.0, origin=(0.0, 0.0)):
    """m x n rectangular cells split into a lower and an upper triangle.

    Follows anuga/abstract_2d_finite_volumes/mesh_factory.py:84-135.
    """
    m = int(m)
    n = int(n)
    delta1 = float(len1) / m
    delta2 = float(len2) / n
    ox, oy = float(origin[0]), float(origin[1])
    points = np.zeros(((m + 1) * (n + 1), 2), dtype=np.float64)
    ii = np.repeat(np.arange(m + 1, dtype=np.float64), n + 1)
    jj = np.tile(np.arange(n + 1, dtype=np.float64), m + 1)
    points[:, 0] = ii * delta1 + ox
    points[:, 1] = jj * delta2 + oy

    ci = np.repeat(np.arange(m, dtype=np.int64), n)
    cj = np.tile(np.arange(n, dtype=np.int64), m)
    i1 = ci * (n + 1) + (cj + 1)
    i2 = ci * (n + 1) + cj
    i3 = (ci + 1) * (n + 1) + (cj + 1)
    i4 = (ci + 1) * (n + 1) + cj
    elements = np.empty((2 * m * n, 3), dtype=np.int64)
    e = elements.reshape(m * n, 2, 3)
    e[:, 0, 0], e[:, 0, 1], e[:, 0, 2] = i4, i3, i2   # lower
    e[:, 1, 0], e[:, 1, 1], e[:, 1, 2] = i1, i2, i3   # upper

    cell = np.arange(m * n, dtype=np.int64)
    boundary = {}
    sel = ci == m - 1
    _boundary_dict(2 * cell[sel], np.full(sel.sum(), 2, np.int64), "right", boundary)
    sel = cj == 0
    _boundary_dict(2 * cell[sel], np.full(sel.sum(), 1, np.int64), "bottom", boundary)
    sel = ci == 0
    _boundary_dict(2 * cell[sel] + 1, np.full(sel.sum(), 2, np.int64), "left", boundary)
    sel = cj == n - 1
    _boundary_dict(2 * cell[sel] + 1, np.full(sel.sum(), 1, np.int64), "top", boundary)
    return points, elements, boundary


# ---------------------------------------------------------------------------
# Connectivity
# ---------------------------------------------------------------------------

def build_neighbour_structure(number_of_nodes, triangles):
    """Neighbour triangle and neighbour edge of every (triangle, edge).

    Edge i of triangle k is the directed segment (v[i+1], v[i+2]); its neighbour
    is the triangle holding the reversed segment (neighbour_table.cpp:63-212).
    Returns (neighbours, neighbour_edges, number_of_boundaries) with -1 where
    there is no neighbour. Raises ValueError on a duplicated directed edge, the
    reference's error code 1.
    """
    t = np.ascontiguousarray(triangles, dtype=np.int64)
    N = t.shape[0]
    nn = np.int64(number_of_nodes)
    a = np.stack([t[:, 1], t[:, 2], t[:, 0]], axis=1).ravel()
    b = np.stack([t[:, 2], t[:, 0], t[:, 1]], axis=1).ravel()
    fwd = a * nn + b
    order = np.argsort(fwd, kind="stable")
    sk = fwd[order]
    if sk.size > 1 and np.any(sk[1:] == sk[:-1]):
        raise ValueError("duplicate directed edge: mesh is not consistently oriented")
    rev = b * nn + a
    pos = np.searchsorted(sk, rev)
    pos_c = np.minimum(pos, sk.size - 1)
    found = (pos < sk.size) & (sk[pos_c] == rev)
    flat = order[pos_c]
    neighbours = np.where(found, flat // 3, -1).reshape(N, 3)
    neighbour_edges = np.where(found, flat % 3, -1).reshape(N, 3)
    number_of_boundaries = 3 - found.reshape(N, 3).sum(axis=1).astype(np.int64)
    return (neighbours.astype(np.int64), neighbour_edges.astype(np.int64),
            number_of_boundaries)


class Mesh:
    """Static arrays of a triangular mesh, in the reference's layout.

    Attributes (N triangles, M boundary edges):
      nodes (Nn,2), triangles (N,3) int64, vertex_coordinates (3N,2),
      edge_midpoint_coordinates (3N,2), centroid_coordinates (N,2),
      areas (N), normals (N,6), edgelengths (N,3), radii (N),
      neighbours (N,3) with boundary edges = -(j+1), neighbour_edges (N,3),
      surrogate_neighbours (N,3), number_of_boundaries (N),
      boundary {(vol,edge): tag}, boundary_cells (M), boundary_edges (M),
      tag_boundary_cells {tag: [j, ...]}.
    """

    def __init__(self, points, triangles, boundary=None, use_inscribed_circle=False):
        self.nodes = np.ascontiguousarray(points, dtype=np.float64)
        self.triangles = np.ascontiguousarray(triangles, dtype=np.int64)
        if self.triangles.ndim != 2 or self.triangles.shape[1] != 3:
            raise ValueError("triangles must be an (N,3) integer array")
        if self.triangles.max() >= self.nodes.shape[0]:
            raise ValueError("Vertex indices reference non-existing coordinate sets")
        self.number_of_triangles = N = self.triangles.shape[0]
        self.number_of_nodes = self.nodes.shape[0]
        self.use_inscribed_circle = use_inscribed_circle
        self._compute_geometry()
        self._compute_connectivity(boundary)

    def __len__(self):
        return self.number_of_triangles

    # general_mesh.py:160-285
    def _compute_geometry(self):
        N = self.number_of_triangles
        V = self.nodes[self.triangles.ravel()]
        self.vertex_coordinates = V
        V0 = V[0::3]
        V1 = V[1::3]
        V2 = V[2::3]
        x0, y0 = V0[:, 0], V0[:, 1]
        x1, y1 = V1[:, 0], V1[:, 1]
        x2, y2 = V2[:, 0], V2[:, 1]

        self.areas = -((x1 * y0 - x0 * y1) + (x2 * y1 - x1 * y2)
                       + (x0 * y2 - x2 * y0)) / 2.0
        if not np.all(self.areas > 0.0):
            bad = np.where(self.areas <= 0.0)[0]
            raise ValueError("Degenerate Triangle(s) %s" % str(bad[:20]))

        normals = np.empty((N, 6), dtype=np.float64)
        lengths = np.empty((N, 3), dtype=np.float64)
        for i, (xa, ya, xb, yb) in enumerate(((x1, y1, x2, y2), (x2, y2, x0, y0),
                                              (x0, y0, x1, y1))):
            xn = xb - xa
            yn = yb - ya
            ln = np.sqrt(xn * xn + yn * yn)
            xn = xn / ln
            yn = yn / ln
            normals[:, 2 * i] = yn
            normals[:, 2 * i + 1] = -xn
            lengths[:, i] = ln
        self.normals = normals
        self.edgelengths = lengths

        cen = np.empty((N, 2), dtype=np.float64)
        cen[:, 0] = (x0 + x1 + x2) / 3
        cen[:, 1] = (y0 + y1 + y2) / 3
        self.centroid_coordinates = cen

        if not self.use_inscribed_circle:
            xm0 = (x1 + x2) / 2
            ym0 = (y1 + y2) / 2
            xm1 = (x2 + x0) / 2
            ym1 = (y2 + y0) / 2
            xm2 = (x0 + x1) / 2
            ym2 = (y0 + y1) / 2
            cx, cy = cen[:, 0], cen[:, 1]
            d0 = np.sqrt((cx - xm0) ** 2 + (cy - ym0) ** 2)
            d1 = np.sqrt((cx - xm1) ** 2 + (cy - ym1) ** 2)
            d2 = np.sqrt((cx - xm2) ** 2 + (cy - ym2) ** 2)
            self.radii = np.minimum(np.minimum(d0, d1), d2)
        else:
            a = np.sqrt((x0 - x1) ** 2 + (y0 - y1) ** 2)
            b = np.sqrt((x1 - x2) ** 2 + (y1 - y2) ** 2)
            c = np.sqrt((x2 - x0) ** 2 + (y2 - y0) ** 2)
            self.radii = 2.0 * self.areas / (a + b + c)

        # general_mesh.py:599-628
        E = np.empty((3 * N, 2), dtype=np.float64)
        E[0::3] = 0.5 * (V1 + V2)
        E[1::3] = 0.5 * (V2 + V0)
        E[2::3] = 0.5 * (V0 + V1)
        self.edge_midpoint_coordinates = E

    def _compute_connectivity(self, boundary):
        N = self.number_of_triangles
        nb, nbe, nob = build_neighbour_structure(self.number_of_nodes, self.triangles)
        self.neighbours = nb
        self.neighbour_edges = nbe
        self.number_of_boundaries = nob
        # neighbour_mesh.py:318-346 (before boundary enumeration overwrites neighbours)
        rng = np.arange(N, dtype=np.int64)[:, None]
        self.surrogate_neighbours = np.where(nb < 0, rng, nb).astype(np.int64)

        # neighbour_mesh_ext.pyx:9-31: complete the dictionary with the default tag
        bdict = {} if boundary is None else dict(boundary)
        for (vol, edge) in bdict:
            if not (0 <= vol < N and 0 <= edge < 3):
                raise ValueError("Segment (%d, %d) does not exist" % (vol, edge))
        kv, ke = np.nonzero(nb < 0)
        if len(bdict) == 0:
            for v, e in zip(kv.tolist(), ke.tolist()):
                bdict[(v, e)] = DEFAULT_BOUNDARY_TAG
        else:
            for v, e in zip(kv.tolist(), ke.tolist()):
                if (v, e) not in bdict:
                    bdict[(v, e)] = DEFAULT_BOUNDARY_TAG
        self.boundary = bdict
        self.boundary_length = M = len(bdict)

        # neighbour_mesh.py:441-505: sorted (vol, edge) keys -> -(j+1)
        keys = np.array(list(bdict.keys()), dtype=np.int64).reshape(-1, 2)
        order = np.lexsort((keys[:, 1], keys[:, 0]))
        keys = keys[order]
        self.boundary_cells = np.ascontiguousarray(keys[:, 0])
        self.boundary_edges = np.ascontiguousarray(keys[:, 1])
        j = np.arange(M, dtype=np.int64)
        self.neighbours[self.boundary_cells, self.boundary_edges] = -(j + 1)
        tags = [bdict[(int(v), int(e))] for v, e in keys.tolist()] if M else []
        self.boundary_tags_by_index = tags
        self.tag_boundary_cells = {}
        for t in sorted(set(tags)):
            self.tag_boundary_cells[t] = []
        for jj, t in enumerate(tags):
            self.tag_boundary_cells[t].append(jj)

    def get_boundary_tags(self):
        return list(self.tag_boundary_cells.keys())

    def get_centroid_coordinates(self):
        return self.centroid_coordinates

    def get_vertex_coordinates(self):
        return self.vertex_coordinates

    def get_edge_midpoint_coordinates(self):
        return self.edge_midpoint_coordinates

    def find_triangles(self, points, tol=1.0e-12):
        """Index of a triangle containing each point (barycentric test with a
        relative tolerance; the lowest index on shared edges), -1 outside."""
        pts = np.atleast_2d(np.asarray(points, np.float64))
        V = self.vertex_coordinates.reshape(-1, 3, 2)
        a, b, c = V[:, 0], V[:, 1], V[:, 2]
        det = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (c[:, 0] - a[:, 0]) * (b[:, 1] - a[:, 1])
        out = np.full(len(pts), -1, np.int64)
        for i, (x, y) in enumerate(pts):
            l1 = ((x - a[:, 0]) * (c[:, 1] - a[:, 1]) - (c[:, 0] - a[:, 0]) * (y - a[:, 1])) / det
            l2 = ((b[:, 0] - a[:, 0]) * (y - a[:, 1]) - (x - a[:, 0]) * (b[:, 1] - a[:, 1])) / det
            inside = (l1 >= -tol) & (l2 >= -tol) & (l1 + l2 <= 1.0 + tol)
            hit = np.nonzero(inside)[0]
            if len(hit):
                out[i] = hit[0]
        return out


def graded_square(length, fine_area, coarse_area, fine_box, jitter=0.15, seed=0,
                  origin=(0.0, 0.0)):
    """An unstructured mesh of a square: near-equilateral triangles of about
    `fine_area` m^2 inside `fine_box` = (x0, y0, x1, y1) and `coarse_area` m^2
    elsewhere, like create_domain_from_regions (anuga/extras.py:59,
    anuga/pmesh/mesh_interface.py:28) with one interior region and two
    maximum areas. That builder needs the `triangle` library; this one
    places jittered hexagonal point sets of the two densities (numpy
    default_rng(seed), jitter as a fraction of the spacing), points along the
    four sides, and takes their Delaunay triangulation (scipy). Boundary tags
    'left', 'right', 'bottom', 'top' as rectangular_cross. Returns (points,
    triangles, boundary)."""
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    ox, oy = origin
    L = float(length)

    def lattice(area, box, keep):
        a = np.sqrt(4.0 * area / np.sqrt(3.0))       # equilateral side of that area
        h = a * np.sqrt(3.0) / 2.0
        x0, y0, x1, y1 = box
        ys = np.arange(y0 + h / 2, y1, h)
        rows = []
        for r, y in enumerate(ys):
            xs = np.arange(x0 + (a / 2 if r % 2 else 0.0) + a / 4, x1, a)
            rows.append(np.stack([xs, np.full_like(xs, y)], axis=1))
        p = np.concatenate(rows) if rows else np.empty((0, 2))
        p = p + rng.uniform(-jitter, jitter, p.shape) * a
        return p[keep(p, a)], a

    fx0, fy0, fx1, fy1 = fine_box
    inside = lambda p, m: ((p[:, 0] > fx0 - m) & (p[:, 0] < fx1 + m) &
                           (p[:, 1] > fy0 - m) & (p[:, 1] < fy1 + m))
    fine, af = lattice(fine_area, fine_box, lambda p, a: inside(p, 0.0))
    coarse, ac = lattice(coarse_area, (0.0, 0.0, L, L),
                         lambda p, a: ~inside(p, 0.5 * a) & (p[:, 0] > 0.5 * a) &
                         (p[:, 0] < L - 0.5 * a) & (p[:, 1] > 0.5 * a) & (p[:, 1] < L - 0.5 * a))
    nb = max(int(np.ceil(L / ac)), 1)
    t = np.linspace(0.0, L, nb + 1)
    side = np.concatenate([np.stack([t, np.zeros_like(t)], 1), np.stack([t, np.full_like(t, L)], 1),
                           np.stack([np.zeros_like(t[1:-1]), t[1:-1]], 1),
                           np.stack([np.full_like(t[1:-1], L), t[1:-1]], 1)])
    pts = np.concatenate([side, coarse, fine])
    tri = Delaunay(pts).simplices.astype(np.int64)
    # counter-clockwise, as the area formula of Mesh expects
    P = pts[tri]
    cross = ((P[:, 1, 0] - P[:, 0, 0]) * (P[:, 2, 1] - P[:, 0, 1]) -
             (P[:, 1, 1] - P[:, 0, 1]) * (P[:, 2, 0] - P[:, 0, 0]))
    flip = cross < 0
    tri[flip] = tri[flip][:, [0, 2, 1]]
    tri = tri[np.abs(cross) > 1e-9 * L * L / len(tri)]
    # boundary edges (edge i opposite vertex i) and their sides
    e = np.concatenate([tri[:, [1, 2]], tri[:, [2, 0]], tri[:, [0, 1]]])
    key = np.sort(e, axis=1)
    _, inv, cnt = np.unique(key[:, 0] * len(pts) + key[:, 1], return_inverse=True,
                            return_counts=True)
    single = np.nonzero(cnt[inv] == 1)[0]
    vol, edge = single % len(tri), single // len(tri)
    mid = 0.5 * (pts[e[single, 0]] + pts[e[single, 1]])
    tol = 1e-9 * L
    names = np.where(np.abs(mid[:, 0]) <= tol, "left",
                     np.where(np.abs(mid[:, 0] - L) <= tol, "right",
                              np.where(np.abs(mid[:, 1]) <= tol, "bottom", "top")))
    boundary = {(int(v), int(i)): str(n) for v, i, n in zip(vol, edge, names)}
    pts = pts + np.array([ox, oy])
    return pts, tri, boundary

"""Boundaries read from files: File_boundary and Field_boundary.

The values of a boundary edge come from a NetCDF time series, interpolated
linearly in time and, for an SWW file, in space at the edge's midpoint; the
domain evaluates them on the host at each substep time and uploads them to
the boundary-value slot the flux kernel reads (the mechanism of the other
time-dependent boundaries).

Restated from the reference:
* File_boundary     anuga/abstract_2d_finite_volumes/generic_boundary_conditions.py:518-680
* Field_boundary    anuga/shallow_water/boundaries.py:993-1075
* file_function     anuga/abstract_2d_finite_volumes/file_function.py:29-225
* get_netcdf_file_function  file_function.py:226-497 (sww and tms files)
* Interpolation_function    anuga/fit_interpolate/interpolate.py:710-1120
  (precomputed values per stored time, linear in time)
* the interpolation weights of a point in a triangle: calculate_sigma and the
  inward edge normals of new_triangle (anuga/utilities/quad_tree.c:24-92)

Files are read with scipy's NetCDF-3 reader (the reference writes SWW and
TMS files as NetCDF classic / 64-bit offset). STS files with a boundary
polygon and CSV files are not supported.
"""

from __future__ import annotations

import os

import numpy as np

from .boundaries import KIND_DIRICHLET, Boundary


class Modeltime_too_early(Exception):
    pass


class Modeltime_too_late(Exception):
    pass


# --- spatial interpolation --------------------------------------------------------
def _inward_normal(xa, ya, xb, yb, xo, yo):
    """new_triangle's edge normal (quad_tree.c:24-60): perpendicular to the
    edge a->b, unit length, flipped so that it points away from the opposite
    vertex o's side test, as the reference forms it."""
    ny = xb - xa
    nx = -(yb - ya)
    dot = nx * nx + ny * ny
    nx = nx / np.sqrt(dot)
    ny = ny / np.sqrt(dot)
    flip = (nx * (xo - xb) + ny * (yo - yb)) > 0
    return np.where(flip, -nx, nx), np.where(flip, -ny, ny)


def _sigmas(tri_xy, x, y):
    """calculate_sigma (quad_tree.c:82-92): the three interpolation weights of
    points (x, y) in triangles tri_xy = (x1, y1, x2, y2, x3, y3)."""
    x1, y1, x2, y2, x3, y3 = tri_xy
    nx3, ny3 = _inward_normal(x1, y1, x2, y2, x3, y3)
    nx1, ny1 = _inward_normal(x2, y2, x3, y3, x1, y1)
    nx2, ny2 = _inward_normal(x3, y3, x1, y1, x2, y2)
    s0 = ((x - x2) * nx1 + (y - y2) * ny1) / ((x1 - x2) * nx1 + (y1 - y2) * ny1)
    s1 = ((x - x3) * nx2 + (y - y3) * ny2) / ((x2 - x3) * nx2 + (y2 - y3) * ny2)
    s2 = ((x - x1) * nx3 + (y - y1) * ny3) / ((x3 - x1) * nx3 + (y3 - y1) * ny3)
    return s0, s1, s2


def _locate(nodes, triangles, points, tol=1e-12):
    """For each point, a triangle that contains it (on an edge or vertex: the
    lowest-numbered) and its weights; -1 where no triangle does."""
    tri = np.asarray(triangles, np.int64)
    V = np.asarray(nodes, np.float64)
    xy = (V[tri[:, 0], 0], V[tri[:, 0], 1], V[tri[:, 1], 0], V[tri[:, 1], 1],
          V[tri[:, 2], 0], V[tri[:, 2], 1])
    xmin = np.minimum(np.minimum(xy[0], xy[2]), xy[4])
    xmax = np.maximum(np.maximum(xy[0], xy[2]), xy[4])
    ymin = np.minimum(np.minimum(xy[1], xy[3]), xy[5])
    ymax = np.maximum(np.maximum(xy[1], xy[3]), xy[5])
    found = np.full(len(points), -1, np.int64)
    sig = np.zeros((len(points), 3))
    for p, (x, y) in enumerate(points):
        cand = np.nonzero((xmin <= x) & (x <= xmax) & (ymin <= y) & (y <= ymax))[0]
        if len(cand) == 0:
            continue
        s = _sigmas(tuple(a[cand] for a in xy), x, y)
        inside = (s[0] >= -tol) & (s[1] >= -tol) & (s[2] >= -tol)
        hit = np.nonzero(inside)[0]
        if len(hit):
            h = hit[0]
            found[p] = cand[h]
            sig[p] = (s[0][h], s[1][h], s[2][h])
    return found, sig


class Interpolation_function:
    """Callable f(t, point_id) over a time series stored at times `time`
    (interpolate.py:710-1120): values at the interpolation points are
    precomputed per stored time (spatially, from the mesh's vertex values
    with each point's triangle weights), then interpolated linearly in time,
    Q0 + ratio (Q1 - Q0)."""

    def __init__(self, time, quantities, quantity_names, vertex_coordinates=None,
                 triangles=None, interpolation_points=None, time_thinning=1):
        time = np.asarray(time, np.float64)
        if not np.all(time[1:] - time[:-1] >= 0):
            raise Exception("Time must be a monotonuosly increasing sequence %s" % time)
        self.spatial = vertex_coordinates is not None
        self.time = np.array(time[::time_thinning])
        self.quantity_names = list(quantity_names)
        self.index = 0
        self.interpolation_points = None
        self.precomputed_values = {}
        q = {}
        for name in self.quantity_names:
            a = np.asarray(quantities[name], np.float64)
            q[name] = np.array(a[::time_thinning, :]) if a.ndim == 2 else a
        if interpolation_points is not None and self.spatial:
            if triangles is None:
                raise Exception("Sww file function requires both triangles and "
                                "vertex_coordinates. sts file file function requires the "
                                "latter.")
            pts = np.asarray(interpolation_points, np.float64)
            self.interpolation_points = pts
            tri = np.asarray(triangles, np.int64)
            k, sig = _locate(vertex_coordinates, tri, pts)
            self.indices_outside_mesh = np.nonzero(k < 0)[0]
            # one row of the interpolation matrix per point: its triangle's
            # vertices in increasing order, the order the sparse product adds
            cols = np.zeros((len(pts), 3), np.int64)
            wts = np.zeros((len(pts), 3))
            for p in range(len(pts)):
                if k[p] < 0:
                    continue
                order = np.argsort(tri[k[p]], kind="stable")
                cols[p] = tri[k[p]][order]
                wts[p] = sig[p][order]
            m, nt = len(pts), len(self.time)
            for name in self.quantity_names:
                Q = q[name]
                out = np.zeros((nt, m))
                for i in range(nt):
                    row = Q[i] if Q.ndim == 2 else Q
                    v = wts[:, 0] * row[cols[:, 0]]
                    v = v + wts[:, 1] * row[cols[:, 1]]
                    v = v + wts[:, 2] * row[cols[:, 2]]
                    v[k < 0] = 0.0
                    out[i] = v
                self.precomputed_values[name] = out
        else:
            for name in self.quantity_names:
                self.precomputed_values[name] = q[name]

    def _bracket(self, t):
        if t < self.time[0]:
            raise Modeltime_too_early("Model time %.16f is not contained in function domain "
                                      "[%.16f:%.16f].\n" % (t, self.time[0], self.time[-1]))
        if t > self.time[-1]:
            raise Modeltime_too_late("Model time %.16f is not contained in function domain "
                                     "[%.16f:%.16f].\n" % (t, self.time[0], self.time[-1]))
        while t > self.time[self.index]:
            self.index += 1
        while t < self.time[self.index]:
            self.index -= 1
        if t == self.time[self.index]:
            return 0.0
        return (t - self.time[self.index]) / (self.time[self.index + 1] - self.time[self.index])

    def values(self, t, point_ids=None):
        """(len(point_ids), quantities) at time t (all points of a spatial
        function; the one series of a time-only one)."""
        ratio = self._bracket(t)
        cols = []
        for name in self.quantity_names:
            Q = self.precomputed_values[name]
            if self.spatial:
                Q0 = Q[self.index] if point_ids is None else Q[self.index, point_ids]
                if ratio > 0:
                    Q1 = Q[self.index + 1] if point_ids is None else Q[self.index + 1, point_ids]
                    cols.append(Q0 + ratio * (Q1 - Q0))
                else:
                    cols.append(np.array(Q0, np.float64))
            else:
                Q0 = Q[self.index]
                cols.append(Q0 + ratio * (Q[self.index + 1] - Q0) if ratio > 0 else Q0)
        return np.stack([np.atleast_1d(c) for c in cols], axis=-1)

    def __call__(self, t, point_id=None, x=None, y=None):
        if self.spatial and point_id is None:
            raise Exception("Either point_id or x and y must be specified")
        v = self.values(t, None if not self.spatial else [point_id])
        return v[0]


def _read_netcdf(filename):
    from scipy.io import netcdf_file
    with netcdf_file(filename, "r", mmap=False) as f:
        variables = {k: np.array(v[:]) for k, v in f.variables.items()}
        attrs = {k: (v.decode() if isinstance(v, bytes) else v)
                 for k, v in f._attributes.items()}
    return variables, attrs


def get_netcdf_file_function(filename, quantity_names=None, interpolation_points=None,
                             domain_starttime=None, use_relative_time=False, time_thinning=1,
                             time_limit=None, verbose=False, boundary_polygon=None):
    """file_function.py:226-497 for sww and tms files: (Interpolation_function,
    the file's starttime)."""
    if isinstance(quantity_names, str):
        quantity_names = [quantity_names]
    if quantity_names is None or len(quantity_names) < 1:
        raise Exception("No quantities are specified in file_function")
    if boundary_polygon is not None:
        raise NotImplementedError("boundary_polygon (sts files) is not supported")
    var, attrs = _read_netcdf(filename)
    if interpolation_points is not None:
        interpolation_points = np.array(interpolation_points, np.float64)
        if interpolation_points.shape[1] != 2:
            raise AssertionError("Points must by N x 2. I got %d" % interpolation_points.shape[1])
    missing = [q for q in ["time"] + list(quantity_names) if q not in var]
    if missing:
        raise Exception("Quantities %s could not be found in file %s" % (str(missing), filename))
    spatial = "x" in var and "y" in var
    if filename.endswith("tms") and spatial:
        raise Exception("Files of type TMS must not contain spatial information")
    if filename.endswith("sww") and not spatial:
        raise Exception("Files of type SWW must contain spatial information")
    try:
        starttime = float(attrs["starttime"])
    except (KeyError, ValueError):
        raise Exception("Could not read starttime from file %s" % filename)
    time = np.asarray(var["time"], np.float64)
    if not use_relative_time:
        time = time + starttime
    upper = len(time)
    if upper == 0:
        raise AssertionError("Time vector obtained from file %s has length 0" % filename)
    if time_limit is not None:
        time_limit = time_limit - starttime
        for i, t in enumerate(time):
            if t > time_limit:
                upper = i
                break
        if upper == 0:
            raise AssertionError("Time vector is zero. Requested time limit is %f" % time_limit)
    time = time[:upper]
    vertex_coordinates = triangles = None
    if spatial:
        xll = float(attrs.get("xllcorner", 0.0))
        yll = float(attrs.get("yllcorner", 0.0))
        vertex_coordinates = np.stack([np.asarray(var["x"], np.float64),
                                       np.asarray(var["y"], np.float64)], axis=1)
        if filename.endswith("sww"):
            triangles = np.asarray(var["volumes"], np.int64)
        if interpolation_points is not None:
            interpolation_points[:, 0] -= xll
            interpolation_points[:, 1] -= yll
    else:
        interpolation_points = None
    if domain_starttime is not None and domain_starttime > starttime:
        time = time - domain_starttime + starttime
    quantities = {name: np.asarray(var[name], np.float64) for name in quantity_names}
    if spatial:
        for name in quantity_names:
            if quantities[name].ndim == 2:
                quantities[name] = quantities[name][:upper]
    else:
        for name in quantity_names:
            quantities[name] = quantities[name][:upper]
    f = Interpolation_function(time, quantities, quantity_names, vertex_coordinates, triangles,
                               interpolation_points, time_thinning=time_thinning)
    return f, starttime


def file_function(filename, domain=None, quantities=None, interpolation_points=None,
                  use_relative_time=True, time_thinning=1, time_limit=None, verbose=False,
                  use_cache=False, boundary_polygon=None, output_centroids=False):
    """file_function.py:29-168: a callable over an sww or tms file; a later
    file starttime moves the domain's starttime to it."""
    if quantities is None:
        quantities = ["stage", "xmomentum", "ymomentum"]
    ext = os.path.splitext(filename)[1]
    if ext not in (".csv", ".sww", ".tms", ".sts"):
        raise AssertionError("Extension should be csv  sww, tms or sts ")
    if ext not in (".sww", ".tms"):
        raise NotImplementedError("file_function: %s files are not supported" % ext)
    domain_starttime = domain.starttime if domain is not None else None
    f, starttime = get_netcdf_file_function(filename, quantities, interpolation_points,
                                            domain_starttime,
                                            use_relative_time=use_relative_time,
                                            time_thinning=time_thinning, time_limit=time_limit,
                                            verbose=verbose, boundary_polygon=boundary_polygon)
    f.starttime = starttime
    f.filename = filename
    if domain is not None and starttime > domain.starttime:
        domain.set_starttime(starttime)
    return f


class File_boundary(Boundary):
    """Values from an sww or tms file at the midpoint of each boundary edge,
    linear in time (generic_boundary_conditions.py:518-680). A model time
    past the file's end raises, or passes to default_boundary when given."""
    kind = "file"
    code = KIND_DIRICHLET
    time_dependent = True

    def __init__(self, filename, domain, time_thinning=1, time_limit=None,
                 boundary_polygon=None, default_boundary=None, use_cache=False,
                 verbose=False):
        mesh = domain.mesh
        keys = sorted(zip(mesh.boundary_cells.tolist(), mesh.boundary_edges.tolist()))
        em = mesh.edge_midpoint_coordinates.reshape(-1, 3, 2)
        xll, yll = _domain_origin(domain)
        self.midpoint_coordinates = np.array([[em[v, e, 0] + xll, em[v, e, 1] + yll]
                                              for v, e in keys], np.float64)
        self.boundary_indices = {key: i for i, key in enumerate(keys)}
        # boundary edge j of the mesh -> its interpolation point
        self._point_of_edge = np.array([self.boundary_indices[(int(v), int(e))] for v, e in
                                        zip(mesh.boundary_cells, mesh.boundary_edges)],
                                       np.int64)
        self.F = file_function(filename, domain, quantities=["stage", "xmomentum", "ymomentum"],
                               interpolation_points=self.midpoint_coordinates,
                               time_thinning=time_thinning, time_limit=time_limit,
                               use_cache=use_cache, verbose=verbose,
                               boundary_polygon=boundary_polygon)
        if not (default_boundary is None or isinstance(default_boundary, Boundary)):
            raise AssertionError("Keyword argument default_boundary must be either None or a "
                                 "boundary object.\n I got %s" % str(default_boundary))
        self.default_boundary = default_boundary
        self.default_boundary_invoked = False
        self.domain = domain
        self.verbose = verbose
        q = self.F.values(0.0 if self.F.time[0] <= 0.0 <= self.F.time[-1] else self.F.time[0],
                          [0] if self.F.spatial else None)
        if q.shape[-1] != 3:
            raise AssertionError("Values specified in file %s must be  a list or an array of "
                                 "length 3" % filename)

    def __repr__(self):
        return "File boundary"

    def values_at(self, t, ids):
        """(len(ids), 3) boundary values of the mesh's boundary edges ids at
        model time t."""
        try:
            if self.F.spatial:
                return self.F.values(t, self._point_of_edge[ids])
            return np.repeat(self.F.values(t).reshape(1, 3), len(ids), axis=0)
        except Modeltime_too_late:
            if self.default_boundary is None:
                raise
            self.default_boundary_invoked = True
            d = self.default_boundary
            if getattr(d, "time_dependent", False):
                return _values_at(d, t, ids)
            if getattr(d, "code", None) == KIND_DIRICHLET:
                return np.repeat(np.asarray(d.dirichlet_values[:3]).reshape(1, 3), len(ids), 0)
            raise NotImplementedError("default_boundary must be a Dirichlet or time boundary")

    def values(self, t):
        return self.values_at(t, np.arange(self.domain.boundary_length))


class Field_boundary(Boundary):
    """File_boundary with mean_stage added to the stage
    (boundaries.py:993-1075)."""
    kind = "field"
    code = KIND_DIRICHLET
    time_dependent = True

    def __init__(self, filename, domain, mean_stage=0.0, time_thinning=1, time_limit=None,
                 boundary_polygon=None, default_boundary=None, use_cache=False,
                 verbose=False):
        self.file_boundary = File_boundary(filename, domain, time_thinning=time_thinning,
                                           time_limit=time_limit,
                                           boundary_polygon=boundary_polygon,
                                           default_boundary=default_boundary,
                                           use_cache=use_cache, verbose=verbose)
        self.F = self.file_boundary.F
        self.domain = self.file_boundary.domain
        self.mean_stage = mean_stage

    def __repr__(self):
        return "Field boundary"

    def values_at(self, t, ids):
        q = np.array(self.file_boundary.values_at(t, ids), np.float64)
        q[:, 0] += self.mean_stage
        return q

    def values(self, t):
        return self.values_at(t, np.arange(self.domain.boundary_length))


def _domain_origin(domain):
    geo = getattr(domain, "geo_reference", None)
    if geo is None:
        return 0.0, 0.0
    return float(geo.get_xllcorner()), float(geo.get_yllcorner())


def _values_at(B, t, ids):
    """Per-edge values of a time-dependent boundary: its own per-edge
    values, or its one value on every edge."""
    if hasattr(B, "values_at"):
        return B.values_at(t, ids)
    return np.repeat(np.asarray(B.values(t), np.float64).reshape(1, 3), len(ids), axis=0)

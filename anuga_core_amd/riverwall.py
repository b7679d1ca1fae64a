"""Riverwalls: thin walls along mesh edges whose discharge follows a weir law.

Host-side set-up of the riverwall data the device flux kernel reads, with the
interface of the reference's RiverWall (anuga/structures/riverwall.py):
domain.riverwallData.create_riverwalls(riverwalls, riverwallPar,
default_riverwallPar, tol, verbose, output_dir), also bound as
domain.create_riverwalls.

* Edges whose midpoint lies within `tol` of a wall segment are flagged
  (edge_flux_type = 1) and take the crest elevation interpolated linearly
  along the segment (riverwall.py:310-380); later segments and walls
  overwrite earlier ones.
* The flagged half-edges, in (triangle, edge) order, index the elevation
  and the hydraulic-table row (riverwall.py:382-417).
* The hydraulic table has one row per wall, columns Qfactor, s1, s2, h1, h2,
  from riverwallPar[name] or the defaults (riverwall.py:419-469).

The flux treatment itself (the wall as a bed step of the crest height between
the two half-edges, and the Villemonte weir blend) runs on the device.
"""

from __future__ import annotations

import os

import numpy as np

MAX_FLOAT = 1.0e36          # anuga/config.py max_float (placeholder elevation)


class RiverWall:
    # the column order is the device record's (and the reference C's)
    hydraulic_variable_names = ("Qfactor", "s1", "s2", "h1", "h2")

    def __init__(self, domain):
        self.domain = domain
        self.ncol_hydraulic_properties = len(self.hydraulic_variable_names)
        self.default_riverwallPar = {"Qfactor": 1.0, "s1": 0.9, "s2": 0.95, "h1": 1.0, "h2": 1.5}
        self.names = []
        self.riverwall_elevation = np.array([-9.0e20])
        self.hydraulic_properties_rowIndex = np.array([-1_000_000_000])
        self.riverwall_edges = np.array([-1_000_000_000])
        self.hydraulic_properties = np.array([[-9.0e20]])
        self.input_riverwall_geo = None
        self.input_riverwallPar = None

    # -- set-up ------------------------------------------------------------------
    def create_riverwalls(self, riverwalls, riverwallPar={}, default_riverwallPar={},
                          tol=1.0e-4, verbose=True, output_dir=None):
        """Flag the mesh edges on the named 3D polylines `riverwalls`
        ({name: [[x, y, z], ...]}) as riverwalls and upload them.

        riverwallPar: {name: {parameter: value}}; default_riverwallPar
        overrides the defaults (Qfactor 1.0, s1 0.9, s2 0.95, h1 1.0, h2 1.5).
        Existing riverwall data is replaced.
        """
        domain = self.domain
        if not domain.get_using_discontinuous_elevation():
            raise Exception("Riverwalls are currently only supported for discontinuous "
                            "elevation flow algorithms")
        if len(self.names) > 0:
            if verbose:
                print("Warning: There seems to be existing riverwall data")
                print("It will be deleted and overwritten with this function call")
            self.__init__(domain)
        self.input_riverwall_geo = riverwalls
        self.input_riverwallPar = riverwallPar

        for key in default_riverwallPar:
            if key not in self.default_riverwallPar:
                raise Exception("Key %s in default_riverwallPar not recognized" % key)
        defaults = dict(self.default_riverwallPar)
        defaults.update(default_riverwallPar)
        self.default_riverwallPar = defaults
        for name, par in riverwallPar.items():
            if name not in riverwalls:
                raise Exception("Key %s in riverwallPar has no corresponding key in riverwall"
                                % name)
            for key in par:
                if key not in defaults:
                    raise Exception("Hydraulic parameter named %s not recognised in "
                                    "default_riverwallPar" % key)

        names = list(riverwalls.keys())
        exy = domain.mesh.edge_midpoint_coordinates
        llx, lly = _lower_left(domain)
        nedge = exy.shape[0]
        flux_type = np.zeros(nedge, np.int64)
        elevation = np.zeros(nedge) - MAX_FLOAT
        row = np.zeros(nedge) - 1.0
        for i, name in enumerate(names):
            wall = riverwalls[name]
            for j in range(len(wall) - 1):
                start, end = wall[j], wall[j + 1]
                if len(start) != 3 or len(end) != 3:
                    raise Exception("Each riverwall coordinate must have at exactly 3 values "
                                    "[xyz]")
                seg_len = ((start[0] - end[0]) ** 2 + (start[1] - end[1]) ** 2) ** 0.5
                if seg_len < tol:
                    if verbose:
                        print("  Segment with length < tolerance %s ignored" % tol)
                    continue
                # unit vector along the segment; along-segment coordinate and
                # squared normal distance of every edge midpoint
                ux = -(start[0] - end[0]) / seg_len
                uy = -(start[1] - end[1]) / seg_len
                px = exy[:, 0] - (start[0] - llx)
                py = exy[:, 1] - (start[1] - lly)
                dist = (px ** 2 + py ** 2) ** 0.5
                along = px * ux + py * uy
                normal_sq = dist ** 2.0 - along ** 2.0
                on = ((normal_sq < tol ** 2) & (along > 0.0 - tol)
                      & (along < seg_len + tol)).nonzero()[0]
                if len(on) == 0:
                    continue
                flux_type[on] = 1
                w0 = np.clip(along[on] / seg_len, 0.0, 1.0)
                elevation[on] = start[2] * (1.0 - w0) + w0 * end[2]
                row[on] = i

        edges = (flux_type == 1).nonzero()[0]
        self.names = names
        self.riverwall_edges = edges
        self.riverwall_elevation = elevation[edges]
        self.hydraulic_properties_rowIndex = row[edges].astype(np.int64)

        table = np.full((len(names), self.ncol_hydraulic_properties), np.nan)
        for i, name in enumerate(names):
            par = riverwallPar.get(name, {})
            for j, var in enumerate(self.hydraulic_variable_names):
                table[i, j] = par[var] if var in par else defaults[var]
        for i, name in enumerate(names):
            if table[i, 1] >= table[i, 2]:
                raise Exception("s1 >= s2 on riverwall %s. This is not allowed" % name)
            if table[i, 1] < 0.0 or table[i, 2] < 0.0:
                raise Exception("s1 and s2 must be positive, with s1<s2")
            if table[i, 3] >= table[i, 4]:
                raise Exception("h1 >= h2 on riverwall %s. This is not allowed" % name)
            if table[i, 3] < 0.0 or table[i, 4] < 0.0:
                raise Exception("h1 and h2 must be positive, with h1<h2")
        self.hydraulic_properties = table

        domain._set_riverwalls(flux_type, self.riverwall_elevation,
                               self.hydraulic_properties_rowIndex, table)

        info, connected = self.check_riverwall_connectedness(verbose=verbose)
        self.export_riverwalls_to_text(output_dir=output_dir)
        if verbose:
            print(info)
            if not connected:
                raise Exception("Riverwall discontinuity -- possible round-off error in"
                                "finding edges on wall -- try increasing value of tol")

    # -- queries -------------------------------------------------------------------
    def get_centroids_corresponding_to_edgeInds(self, riverwalledgeInds):
        return (np.asarray(riverwalledgeInds) // 3).astype(int)

    def get_vertices_corresponding_to_edgeInds(self, riverwalledgeInds, checkCoords=True):
        """Flat vertex indices (3k + j) of the two end points of each edge
        3k + i: vertices i+1 and i+2 (mod 3) of triangle k."""
        e = np.asarray(riverwalledgeInds, np.int64)
        k, i = e // 3, e % 3
        v1 = 3 * k + (i + 1) % 3
        v2 = 3 * k + (i + 2) % 3
        if checkCoords:
            V = self.domain.mesh.vertex_coordinates
            E = self.domain.mesh.edge_midpoint_coordinates
            assert np.allclose(E[e], 0.5 * (V[v1] + V[v2]))
        return v1, v2

    def is_vertex_on_boundary(self, vertexIndices):
        """1 where the vertex is an end point of a boundary edge, else 0."""
        mesh = self.domain.mesh
        V = mesh.vertex_coordinates
        bedges = np.array([3 * k + i for (k, i) in mesh.boundary.keys()], np.int64)
        b1, b2 = self.get_vertices_corresponding_to_edgeInds(bedges, checkCoords=False)
        on = {tuple(p) for p in V[np.concatenate([b1, b2])]}
        return np.array([1 if tuple(V[v]) in on else 0 for v in np.asarray(vertexIndices)],
                        np.int64)

    def check_riverwall_connectedness(self, verbose=True):
        """A continuous wall has at most two end points inside the domain
        (points on the domain boundary aside); more suggests edges missed
        through round-off. Returns [message, is_connected]."""
        if len(self.names) == 0:
            return ["  There are no riverwalls\n" if verbose else "", True]
        V = self.domain.mesh.vertex_coordinates
        E = self.domain.mesh.edge_midpoint_coordinates
        info, connected = "", True
        for i, name in enumerate(self.names):
            e = self.riverwall_edges[self.hydraulic_properties_rowIndex == i]
            if len(e) == 0:
                info += ("Riverwall %s was not found on this mesh (if this is wrong, adjust "
                         "tol in create_riverwalls)\n" % name)
                continue
            # each interior wall edge appears twice (once per triangle)
            _, first = np.unique(E[e], axis=0, return_index=True)
            e = e[np.sort(first)]
            v1, v2 = self.get_vertices_corresponding_to_edgeInds(e, checkCoords=False)
            ends = np.concatenate([V[v1], V[v2]])
            on_boundary = np.concatenate([self.is_vertex_on_boundary(v1),
                                          self.is_vertex_on_boundary(v2)])
            _, inverse, counts = np.unique(ends, axis=0, return_inverse=True,
                                           return_counts=True)
            loose = int(((counts[inverse.ravel()] == 1) & (on_boundary == 0)).sum())
            if verbose:
                info += ("  On riverwall %s there are %d endpoints inside the domain "
                         "[ignoring points on the boundary polygon]\n" % (name, loose))
            if loose > 2:
                connected = False
                info += ("  Riverwall %s appears to be discontinuous.\n"
                         "  This suggests there is a gap in the wall, which should not occur\n"
                         % name)
        return [info, connected]

    def export_riverwalls_to_text(self, output_dir=None):
        """One text file per wall in output_dir: its hydraulic parameters, then
        x,y,elevation of each of its edges."""
        if output_dir is None:
            return
        os.makedirs(output_dir, exist_ok=True)
        E = self.domain.mesh.edge_midpoint_coordinates
        llx, lly = _lower_left(self.domain)
        for i, name in enumerate(self.names):
            path = os.path.join(output_dir, os.path.splitext(os.path.basename(name))[0] + ".txt")
            sel = (self.hydraulic_properties_rowIndex == i).nonzero()[0]
            with open(path, "w") as f:
                f.write("## Hydraulic Variable values below ## \n")
                f.write(str(self.hydraulic_variable_names) + "\n")
                f.write(str(self.hydraulic_properties[i, :]) + "\n\n")
                f.write("## xyElevation at edges below. Order may be erratic for parallel runs ##\n")
                for j in sel:
                    e = self.riverwall_edges[j]
                    f.write("%s,%s,%s\n" % (E[e, 0] + llx, E[e, 1] + lly,
                                            self.riverwall_elevation[j]))


def _lower_left(domain):
    geo = getattr(domain.mesh, "geo_reference", None)
    if geo is None:
        return 0.0, 0.0
    return geo.get_xllcorner(), geo.get_yllcorner()

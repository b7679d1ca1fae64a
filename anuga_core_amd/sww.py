"""SWW output: the reference's NetCDF layout for model output, written from
frames that the device prepares.

Layout (anuga/file/sww.py Write_sww.store_header :563-692,
store_triangulation :694-796, write_dynamic_quantities :798-829,
store_static_quantities(_centroid) :875-961, store_quantities(_centroid)
:963-1091; SWW_file :69-435): NetCDF classic with 64-bit offsets (the
reference opens netCDF4 Datasets with format NETCDF3_64BIT), float32
quantities, one record per stored yield. Node values are averages of the
centroid values around each node; where the node depth is below
minimum_storable_height the stored stage is the bed and the momenta are zero.

The file is written by a small NetCDF-3 writer of this module (no netCDF
library is needed): the header and the static data once, then one record per
frame appended at the end of the file with the record count updated in
place, so storing a frame costs one frame of I/O.
"""

from __future__ import annotations

import os
import struct

import numpy as np

MAX_FLOAT = 1.0e36                      # anuga/config.py max_float
DEFAULT_INSTITUTION = "Geoscience Australia"

_NC_DIMENSION, _NC_VARIABLE, _NC_ATTRIBUTE = 0x0A, 0x0B, 0x0C
_TYPES = {np.dtype("int8"): 1, np.dtype("S1"): 2, np.dtype("int16"): 3,
          np.dtype("int32"): 4, np.dtype("float32"): 5, np.dtype("float64"): 6}


def _pad4(n):
    return (n + 3) & ~3


def _name(s):
    b = s.encode("utf-8")
    return struct.pack(">i", len(b)) + b + b"\0" * (_pad4(len(b)) - len(b))


def _attr_bytes(attrs):
    if not attrs:
        return struct.pack(">ii", 0, 0)
    out = [struct.pack(">ii", _NC_ATTRIBUTE, len(attrs))]
    for k, v in attrs.items():
        if isinstance(v, str):
            b = v.encode("utf-8")
            out.append(_name(k) + struct.pack(">ii", 2, len(b)) + b + b"\0" * (_pad4(len(b)) - len(b)))
            continue
        a = np.atleast_1d(np.asarray(v))
        if a.dtype.kind in "iu":
            a = a.astype(">i4")
            t = 4
        else:
            a = a.astype(">f8")
            t = 6
        raw = a.tobytes()
        out.append(_name(k) + struct.pack(">ii", t, a.size) + raw + b"\0" * (_pad4(len(raw)) - len(raw)))
    return b"".join(out)


class NC3Writer:
    """Minimal writer of NetCDF classic files with 64-bit offsets (CDF-2):
    dimensions (at most one unlimited), global and per-variable attributes,
    fixed-size variables written once and record variables appended one
    record at a time."""

    def __init__(self, path):
        self.path = path
        self.dims = {}          # name -> length (None: unlimited)
        self.attrs = {}
        self.vars = {}          # name -> dict(dims, dtype, attrs)
        self._f = None
        self.numrecs = 0

    def create_dimension(self, name, length):
        self.dims[name] = length

    def create_variable(self, name, dtype, dims, attrs=None):
        self.vars[name] = dict(dims=tuple(dims), dtype=np.dtype(dtype), attrs=attrs or {})

    # -- layout ------------------------------------------------------------------
    def _is_record(self, v):
        return len(v["dims"]) > 0 and self.dims[v["dims"][0]] is None

    def _vsize(self, v):
        shape = [self.dims[d] for d in v["dims"] if self.dims[d] is not None]
        return _pad4(int(np.prod(shape, dtype=np.int64)) * v["dtype"].itemsize)

    def _header(self, begins):
        dim_names = list(self.dims)
        out = [b"CDF\x02", struct.pack(">i", self.numrecs)]
        if self.dims:
            out.append(struct.pack(">ii", _NC_DIMENSION, len(self.dims)))
            for n in dim_names:
                out.append(_name(n) + struct.pack(">i", self.dims[n] or 0))
        else:
            out.append(struct.pack(">ii", 0, 0))
        out.append(_attr_bytes(self.attrs))
        if self.vars:
            out.append(struct.pack(">ii", _NC_VARIABLE, len(self.vars)))
            for n, v in self.vars.items():
                out.append(_name(n) + struct.pack(">i", len(v["dims"])))
                out.append(b"".join(struct.pack(">i", dim_names.index(d)) for d in v["dims"]))
                out.append(_attr_bytes(v["attrs"]))
                out.append(struct.pack(">iiq", _TYPES[v["dtype"]], self._vsize(v), begins[n]))
        else:
            out.append(struct.pack(">ii", 0, 0))
        return b"".join(out)

    def end_define(self):
        """Write the header; fixed-size variables follow it, then records."""
        begins = {n: 0 for n in self.vars}
        size = len(self._header(begins))
        pos = size
        for n, v in self.vars.items():
            if not self._is_record(v):
                begins[n] = pos
                pos += self._vsize(v)
        self.record_start = pos
        self.recsize = 0
        for n, v in self.vars.items():
            if self._is_record(v):
                begins[n] = pos + self.recsize
                self.recsize += self._vsize(v)
        self.begins = begins
        header = self._header(begins)
        assert len(header) == size
        self._f = open(self.path, "w+b")
        self._f.write(header)
        self._f.write(b"\0" * (self.record_start - size))

    def write(self, name, data):
        """Data of a fixed-size variable."""
        v = self.vars[name]
        a = np.ascontiguousarray(data, dtype=v["dtype"].newbyteorder(">"))
        shape = tuple(self.dims[d] for d in v["dims"])
        assert a.shape == shape or a.size == int(np.prod(shape)), (name, a.shape, shape)
        self._f.seek(self.begins[name])
        self._f.write(a.tobytes())

    def append_record(self, values, index=None):
        """One record: values[name] for every record variable (index: rewrite
        an existing record)."""
        i = self.numrecs if index is None else index
        buf = bytearray(self.recsize)
        off = 0
        for n, v in self.vars.items():
            if not self._is_record(v):
                continue
            a = np.ascontiguousarray(values[n], dtype=v["dtype"].newbyteorder(">")).tobytes()
            buf[off:off + len(a)] = a
            off += self._vsize(v)
        self._f.seek(self.record_start + i * self.recsize)
        self._f.write(bytes(buf))
        if i >= self.numrecs:
            self.numrecs = i + 1
            self._f.seek(4)
            self._f.write(struct.pack(">i", self.numrecs))
        self._f.flush()

    def set_attribute_array(self, name, data):
        """Overwrite a fixed-size variable in place (range variables)."""
        self.write(name, data)
        self._f.flush()

    def close(self):
        if self._f is not None:
            self._f.close()
            self._f = None


class SWW_file:
    """Writer of one SWW file for a Domain (anuga/file/sww.py SWW_file):
    header and triangulation at construction / store_connectivity, one
    record per store_timestep."""

    def __init__(self, domain):
        self.domain = domain
        self.filename = os.path.join(domain.get_datadir(), domain.get_name() + ".sww")
        self.minimum_storable_height = domain.minimum_storable_height
        self.store_centroids = domain.store_centroids
        self.static_quantities = [q for q, f in domain.quantities_to_be_stored.items() if f == 1]
        self.dynamic_quantities = [q for q, f in domain.quantities_to_be_stored.items() if f == 2]
        if sorted(self.dynamic_quantities) != ["stage", "xmomentum", "ymomentum"]:
            raise Exception("the device frames hold stage, xmomentum and ymomentum; "
                            "quantities_to_be_stored must make exactly these dynamic")
        for q in self.static_quantities:
            if q not in ("elevation", "friction"):
                raise Exception("static SWW quantity %s is not supported" % q)
        self.times = []
        self._ranges = {}
        self._nc = None

    def store_connectivity(self):
        d = self.domain
        mesh = d.mesh
        nodes = mesh.nodes
        N = mesh.number_of_triangles
        P = len(nodes)
        nc = NC3Writer(self.filename)
        geo = getattr(mesh, "geo_reference", None)
        nc.attrs.update(dict(institution=DEFAULT_INSTITUTION,
                             description="Output from anuga.file.sww suitable for plotting",
                             smoothing="Yes", vertices_are_stored_uniquely="False",
                             order=int(d.default_order), revision_number="None",
                             revision_date="None", anuga_version="anuga_core_amd",
                             starttime=float(d.starttime), timezone="UTC"))
        nc.attrs.update(_georef_attributes(geo))
        nc.create_dimension("number_of_volumes", N)
        nc.create_dimension("number_of_triangle_vertices", P)
        nc.create_dimension("number_of_vertices", 3)
        nc.create_dimension("numbers_in_range", 2)
        nc.create_dimension("number_of_points", P)
        nc.create_dimension("number_of_timesteps", None)
        f4 = np.float32
        nc.create_variable("x", f4, ("number_of_points",))
        nc.create_variable("y", f4, ("number_of_points",))
        nc.create_variable("volumes", np.int32, ("number_of_volumes", "number_of_vertices"))
        for q in self.static_quantities:
            nc.create_variable(q, f4, ("number_of_points",))
            nc.create_variable(q + "_range", f4, ("numbers_in_range",))
        if self.store_centroids:
            for q in self.static_quantities:
                nc.create_variable(q + "_c", f4, ("number_of_volumes",))
        for q in self.dynamic_quantities:
            nc.create_variable(q, f4, ("number_of_timesteps", "number_of_points"))
            nc.create_variable(q + "_range", f4, ("numbers_in_range",))
        if self.store_centroids:
            for q in self.dynamic_quantities:
                nc.create_variable(q + "_c", f4, ("number_of_timesteps", "number_of_volumes"))
        nc.create_variable("time", np.float64, ("number_of_timesteps",))
        nc.end_define()
        nc.write("x", nodes[:, 0].astype(f4))
        nc.write("y", nodes[:, 1].astype(f4))
        nc.write("volumes", np.asarray(mesh.triangles, np.int32))
        for q in self.static_quantities:
            node_v, cen_v = d._static_output(q)
            nc.write(q, node_v)
            nc.write(q + "_range", np.array([node_v.min(), node_v.max()], f4))
            if self.store_centroids:
                nc.write(q + "_c", cen_v)
        for q in self.dynamic_quantities:
            self._ranges[q] = np.array([MAX_FLOAT, -MAX_FLOAT], f4)
            nc.write(q + "_range", self._ranges[q])
        self._nc = nc

    def store_timestep(self):
        d = self.domain
        t = d.relative_time
        node_v, cen_v = d._sww_frame(self.minimum_storable_height)
        index = None
        if self.times and t <= self.times[-1]:
            # a repeated time overwrites its record (Write_sww.store_quantities)
            hits = np.nonzero(np.abs(np.asarray(self.times) - t) < 1.0e-14)[0]
            if len(hits):
                index = int(hits[0])
        rec = {"time": np.array([t])}
        names = ("stage", "xmomentum", "ymomentum")
        for i, q in enumerate(names):
            rec[q] = node_v[i]
            if self.store_centroids:
                rec[q + "_c"] = cen_v[i]
            r = self._ranges[q]
            r[0] = min(r[0], node_v[i].min())
            r[1] = max(r[1], node_v[i].max())
            self._nc.set_attribute_array(q + "_range", r)
        self._nc.append_record(rec, index)
        if index is None:
            self.times.append(t)

    def close(self):
        if self._nc is not None:
            self._nc.close()
            self._nc = None


def _georef_attributes(geo):
    """Geo_reference.write_NetCDF (coordinate_transforms/geo_reference.py:169-185)
    with the reference's defaults when the mesh has none."""
    get = lambda n, dflt: getattr(geo, n, dflt) if geo is not None else dflt
    return dict(xllcorner=float(get("xllcorner", 0.0)), yllcorner=float(get("yllcorner", 0.0)),
                zone=int(get("zone", -1)), hemisphere=str(get("hemisphere", "undefined")),
                false_easting=int(get("false_easting", 500000)),
                false_northing=int(get("false_northing", 10000000)),
                datum=str(get("datum", "wgs84")), projection=str(get("projection", "UTM")),
                units=str(get("units", "m")))


def read_sww(filename):
    """Every variable and global attribute of an SWW file, as numpy arrays
    and Python values (scipy's NetCDF-3 reader)."""
    from scipy.io import netcdf_file
    with netcdf_file(filename, "r", mmap=False) as f:
        out = {k: np.array(v[:]) for k, v in f.variables.items()}
        attrs = {k: (v.decode() if isinstance(v, bytes) else v)
                 for k, v in f._attributes.items()}
        dims = dict(f.dimensions)
    return out, attrs, dims


def node_structure(triangles, number_of_nodes):
    """number_of_triangles_per_node and vertex_value_indices of
    General_mesh.build_inverted_triangle_structure (general_mesh.py:790-855):
    counts per node (lone nodes 0) and the flat vertex indices sorted by node
    with numpy's default argsort, whose order within a node is the summation
    order of the node averages."""
    flat = np.asarray(triangles).ravel()
    counts = np.bincount(flat, minlength=number_of_nodes).astype(np.int64)
    vvi = np.argsort(flat).astype(np.int64)
    return counts, vvi

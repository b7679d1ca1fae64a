"""ctypes bindings of the two CPU kernel sets used as the oracle.

TEST INFRASTRUCTURE ONLY (imported by tests, the smoke check and the CPU
baseline of the benchmark; never by the product package).

* `OracleKernels` drives oracle/liboracle.so, the plain-C restatement in
  oracle/de_oracle.c.
* `RefKernels` drives oracle/_ref/libref_sw.so and libref_quantity.so, the
  reference's own sw_domain_orig.c and quantity.c compiled unmodified, through
  a ctypes mirror of its `struct domain` (anuga/shallow_water/sw_domain.h:14-113)
  filled the way the Cython shim fills it
  (anuga/shallow_water/sw_domain_orig_ext.pyx:120-355).

Both expose the same methods on an `OracleDomain`'s arrays:
protect, extrapolate, compute_fluxes(substep), friction, update(dt),
fix_negative, backup, saxpy.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int64)


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _i(a):
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def _load(path):
    if not os.path.exists(path):
        raise FileNotFoundError(
            "%s is not built; run `make -C oracle` (and `make -C oracle ref` "
            "where the reference tree is present)" % path)
    return C.CDLL(path)


# ---------------------------------------------------------------------------
# The C restatement
# ---------------------------------------------------------------------------

class _OMesh(C.Structure):
    _fields_ = [("N", C.c_int64), ("M", C.c_int64),
                ("neighbours", _ip), ("neighbour_edges", _ip), ("surrogate", _ip),
                ("nboundaries", _ip), ("tri_full_flag", _ip),
                ("normals", _dp), ("edgelengths", _dp), ("radii", _dp), ("areas", _dp),
                ("centroids", _dp), ("edge_midpoints", _dp),
                ("edge_flux_type", _ip), ("edge_river_wall_counter", _ip),
                ("riverwall_elevation", _dp), ("riverwall_rowIndex", _ip),
                ("ncol_riverwall_hydraulic_properties", C.c_int64),
                ("riverwall_hydraulic_properties", _dp)]


class _OParams(C.Structure):
    _fields_ = [("g", C.c_double), ("epsilon", C.c_double),
                ("minimum_allowed_height", C.c_double),
                ("beta_w", C.c_double), ("beta_w_dry", C.c_double),
                ("beta_uh", C.c_double), ("beta_uh_dry", C.c_double),
                ("beta_vh", C.c_double), ("beta_vh_dry", C.c_double),
                ("low_froude", C.c_int64),
                ("extrapolate_velocity_second_order", C.c_int64)]


_STATE_FIELDS = ["w", "uh", "vh", "z", "h", "w_e", "uh_e", "vh_e", "z_e", "h_e",
                 "w_v", "uh_v", "vh_v", "z_v", "h_v", "w_b", "uh_b", "vh_b",
                 "ex_w", "ex_uh", "ex_vh", "si_w", "si_uh", "si_vh",
                 "xwork", "ywork", "max_speed", "edge_timestep"]


class _OState(C.Structure):
    _fields_ = [(f, _dp) for f in _STATE_FIELDS]


def _state_arrays(d):
    """Map the OracleDomain arrays onto the restatement's state fields."""
    return dict(w=d.stage_c, uh=d.xmom_c, vh=d.ymom_c, z=d.elev_c, h=d.height_c,
                w_e=d.stage_e, uh_e=d.xmom_e, vh_e=d.ymom_e, z_e=d.elev_e, h_e=d.height_e,
                w_v=d.stage_v, uh_v=d.xmom_v, vh_v=d.ymom_v, z_v=d.elev_v, h_v=d.height_v,
                w_b=d.stage_b, uh_b=d.xmom_b, vh_b=d.ymom_b,
                ex_w=d.stage_ex, ex_uh=d.xmom_ex, ex_vh=d.ymom_ex,
                si_w=d.stage_si, si_uh=d.xmom_si, si_vh=d.ymom_si,
                xwork=d.x_centroid_work, ywork=d.y_centroid_work,
                max_speed=d.max_speed, edge_timestep=d.edge_timestep)


class _OLts(C.Structure):
    _fields_ = [("max_flux_update_frequency", C.c_int64), ("call", C.c_int64),
                ("cyclic_number_of_steps", C.c_int64), ("local_timestep", C.c_double),
                ("allow_timestep_increase", C.c_int64),
                ("already_computed_flux", _ip), ("flux_update_frequency", _ip),
                ("update_next_flux", _ip), ("update_extrapolation", _ip),
                ("edge_flux_work", _dp), ("pressuregrad_work", _dp)]


class OracleKernels:
    """Kernels of oracle/de_oracle.c. `gather=True` uses the per-triangle
    (OpenMP) flux loop instead of the edge-once-and-mirror loop; with local
    flux updating on, the flux call is the edge-once loop with its skip test."""

    def __init__(self, gather=False):
        lib = _load(os.path.join(HERE, "liboracle.so"))
        self.lib = lib
        self.gather = gather
        lib.oracle_pow_7_3.restype = C.c_double
        lib.oracle_pow_7_3.argtypes = [C.c_double]
        lib.oracle_pow.restype = C.c_double
        lib.oracle_pow.argtypes = [C.c_double, C.c_double]
        lib.oracle_adjust_edgeflux_with_weir.argtypes = [_dp] + [C.c_double] * 9 + [_dp]
        lib.oracle_protect.restype = C.c_double
        lib.oracle_extrapolate.restype = C.c_int64
        lib.oracle_compute_fluxes.restype = C.c_double
        lib.oracle_compute_fluxes_gather.restype = C.c_double
        lib.oracle_update.restype = C.c_int64
        lib.oracle_fix_negative_cells.restype = C.c_int64
        lib.oracle_compute_fluxes.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                              C.c_double, _dp, _dp, _dp, _ip]
        lib.oracle_compute_fluxes_gather.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p,
                                                     C.c_int64, C.c_double, _dp, _dp]
        lib.oracle_manning_friction_flat.argtypes = [C.c_int64, C.c_double, C.c_double,
                                                     _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        lib.oracle_update.argtypes = [C.c_int64, C.c_double, _dp, _dp, _dp]
        lib.oracle_manning_friction_sloped.argtypes = [C.c_int64, C.c_double, C.c_double, _dp,
                                                       _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        lib.oracle_fix_negative_cells.argtypes = [C.c_int64, _ip, _dp, _dp, _dp, _dp]
        lib.oracle_saxpy.argtypes = [C.c_int64, C.c_double, C.c_double, _dp, _dp]
        lib.oracle_backup.argtypes = [C.c_int64, _dp, _dp]
        lib.oracle_flux_function_central.argtypes = [
            _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
            C.c_double, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp, C.c_int64]
        lib.oracle_rotate.argtypes = [_dp, C.c_double, C.c_double]
        lib.oracle_compute_fluxes_lts.restype = C.c_double
        lib.oracle_compute_fluxes_lts.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p,
                                                  C.c_double, _dp, C.c_void_p]
        lib.oracle_compute_flux_update_frequency.argtypes = [C.c_void_p, C.c_void_p,
                                                             C.c_double, C.c_void_p]
        self.lts = None

    def set_max_flux_update_frequency(self, M):
        """Local flux updating with the reference's initial arrays
        (shallow_water_domain.py:404-426, generic_domain.py:386) and fresh
        static counters; M == 1 switches it off."""
        if M == 1:
            self.lts = None
            return
        N = self.d.N
        a = dict(acf=np.zeros(3 * N, np.int64), fuf=np.ones(3 * N, np.int64),
                 unf=np.ones(3 * N, np.int64), uex=np.ones(N, np.int64),
                 efw=np.zeros(9 * N), pgw=np.zeros(3 * N))
        self._lts_arrays = a
        self.d.edge_timestep[:] = 1.0e100
        self.lts = _OLts(M, 0, -1, 0.0, 1, _i(a["acf"]), _i(a["fuf"]), _i(a["unf"]),
                         _i(a["uex"]), _d(a["efw"]), _d(a["pgw"]))

    def compute_flux_update_frequency(self, timestep):
        if self.lts is None:
            return
        self.lib.oracle_compute_flux_update_frequency(C.byref(self.mesh), C.byref(self.state),
                                                      timestep, C.byref(self.lts))

    def bind(self, d):
        self.d = d
        m = d.mesh
        self._keep = [d.neighbours, d.neighbour_edges, d.surrogate_neighbours,
                      d.number_of_boundaries, d.tri_full_flag]
        rw = d.number_of_riverwall_edges > 0
        self.mesh = _OMesh(d.N, d.M, _i(d.neighbours), _i(d.neighbour_edges),
                           _i(d.surrogate_neighbours), _i(d.number_of_boundaries),
                           _i(d.tri_full_flag), _d(d.normals), _d(d.edgelengths),
                           _d(d.radii), _d(d.areas), _d(d.centroids), _d(d.edge_midpoints),
                           _i(d.edge_flux_type) if rw else None,
                           _i(d.edge_river_wall_counter), _d(d.riverwall_elevation),
                           _i(d.riverwall_rowIndex), d.ncol_riverwall_hydraulic_properties,
                           _d(d.riverwall_hydraulic_properties))
        self.state = _OState(**{k: _d(v) for k, v in _state_arrays(d).items()})
        self._edge_flux_work = np.zeros(9 * d.N)
        self._pressuregrad_work = np.zeros(3 * d.N)
        self._done = np.zeros(3 * d.N, dtype=np.int64)
        self._bflux_tri = np.zeros(d.N)

    def _params(self):
        d = self.d
        return _OParams(d.g, d.epsilon, d.minimum_allowed_height, d.beta_w, d.beta_w_dry,
                        d.beta_uh, d.beta_uh_dry, d.beta_vh, d.beta_vh_dry, d.low_froude,
                        d.extrapolate_velocity_second_order)

    def protect(self):
        p = self._params()
        return self.lib.oracle_protect(C.byref(self.mesh), C.byref(p), C.byref(self.state))

    def extrapolate(self):
        p = self._params()
        return self.lib.oracle_extrapolate(C.byref(self.mesh), C.byref(p), C.byref(self.state))

    def compute_fluxes(self, substep, timestep):
        p = self._params()
        d = self.d
        if self.lts is not None:
            return self.lib.oracle_compute_fluxes_lts(
                C.byref(self.mesh), C.byref(p), C.byref(self.state), timestep,
                _d(d.boundary_flux_sum), C.byref(self.lts))
        if self.gather:
            return self.lib.oracle_compute_fluxes_gather(
                C.byref(self.mesh), C.byref(p), C.byref(self.state), substep, timestep,
                _d(d.boundary_flux_sum), _d(self._bflux_tri))
        return self.lib.oracle_compute_fluxes(
            C.byref(self.mesh), C.byref(p), C.byref(self.state), substep, timestep,
            _d(d.boundary_flux_sum), _d(self._edge_flux_work), _d(self._pressuregrad_work),
            _i(self._done))

    def friction(self):
        d = self.d
        if d.use_sloped_mannings:
            self.lib.oracle_manning_friction_sloped(d.N, d.g, d.minimum_allowed_height,
                                                    _d(d.vertex_coordinates), _d(d.stage_c),
                                                    _d(d.elev_v), _d(d.xmom_c), _d(d.ymom_c),
                                                    _d(d.friction_c), _d(d.xmom_si),
                                                    _d(d.ymom_si))
            return
        self.lib.oracle_manning_friction_flat(d.N, d.g, d.minimum_allowed_height,
                                              _d(d.stage_c), _d(d.elev_c), _d(d.xmom_c),
                                              _d(d.ymom_c), _d(d.friction_c),
                                              _d(d.xmom_si), _d(d.ymom_si))

    def update(self, dt):
        d = self.d
        for q, ex, si in ((d.stage_c, d.stage_ex, d.stage_si),
                          (d.xmom_c, d.xmom_ex, d.xmom_si),
                          (d.ymom_c, d.ymom_ex, d.ymom_si)):
            if self.lib.oracle_update(d.N, dt, _d(q), _d(ex), _d(si)) != 0:
                raise AssertionError("update: division by zero in semi implicit update")

    def fix_negative(self):
        d = self.d
        return self.lib.oracle_fix_negative_cells(d.N, _i(d.tri_full_flag), _d(d.stage_c),
                                                  _d(d.elev_c), _d(d.xmom_c), _d(d.ymom_c))

    def backup(self):
        d = self.d
        for q, b in ((d.stage_c, d.stage_bk), (d.xmom_c, d.xmom_bk), (d.ymom_c, d.ymom_bk)):
            self.lib.oracle_backup(d.N, _d(q), _d(b))

    def saxpy(self, a, b):
        d = self.d
        for q, bk in ((d.stage_c, d.stage_bk), (d.xmom_c, d.xmom_bk), (d.ymom_c, d.ymom_bk)):
            self.lib.oracle_saxpy(d.N, a, b, _d(q), _d(bk))

    # single-call entry points used by the known-answer tests
    def flux_function_central(self, normal, ql, qr, h_left, h_right, hle, hre,
                              epsilon, ze, g, low_froude=0):
        ql = np.ascontiguousarray(ql, np.float64)
        qr = np.ascontiguousarray(qr, np.float64)
        ef = np.zeros(3)
        ms = np.zeros(1)
        pf = np.zeros(1)
        self.lib.oracle_flux_function_central(_d(ql), _d(qr), h_left, h_right, hle, hre,
                                              float(normal[0]), float(normal[1]), epsilon,
                                              ze, g, _d(ef), _d(ms), _d(pf), low_froude)
        return ef, ms[0], pf[0]

    def rotate(self, q, normal, direction=1):
        r = np.array(q, dtype=np.float64)
        n2 = -normal[1] if direction == -1 else normal[1]
        self.lib.oracle_rotate(_d(r), float(normal[0]), float(n2))
        return r

    def adjust_edgeflux_with_weir(self, edgeflux, h_left, h_right, g, weir_height,
                                  Qfactor, s1, s2, h1, h2, max_speed):
        ef = np.array(edgeflux, dtype=np.float64)
        ms = np.array([max_speed], dtype=np.float64)
        self.lib.oracle_adjust_edgeflux_with_weir(_d(ef), h_left, h_right, g, weir_height,
                                                  Qfactor, s1, s2, h1, h2, _d(ms))
        return ef, ms[0]


# ---------------------------------------------------------------------------
# The reference's own C, through a mirror of its struct domain
# ---------------------------------------------------------------------------

_REF_FIELDS = [
    ("number_of_elements", C.c_int64), ("boundary_length", C.c_int64),
    ("number_of_riverwall_edges", C.c_int64), ("epsilon", C.c_double), ("H0", C.c_double),
    ("g", C.c_double), ("optimise_dry_cells", C.c_int64), ("evolve_max_timestep", C.c_double),
    ("extrapolate_velocity_second_order", C.c_int64), ("minimum_allowed_height", C.c_double),
    ("maximum_allowed_speed", C.c_double), ("low_froude", C.c_int64),
    ("timestep_fluxcalls", C.c_int64),
    ("beta_w", C.c_double), ("beta_w_dry", C.c_double), ("beta_uh", C.c_double),
    ("beta_uh_dry", C.c_double), ("beta_vh", C.c_double), ("beta_vh_dry", C.c_double),
    ("max_flux_update_frequency", C.c_int64), ("ncol_riverwall_hydraulic_properties", C.c_int64),
    ("neighbours", _ip), ("neighbour_edges", _ip), ("surrogate_neighbours", _ip),
    ("normals", _dp), ("edgelengths", _dp), ("radii", _dp), ("areas", _dp),
    ("edge_flux_type", _ip), ("tri_full_flag", _ip), ("already_computed_flux", _ip),
    ("max_speed", _dp), ("vertex_coordinates", _dp), ("edge_coordinates", _dp),
    ("centroid_coordinates", _dp), ("number_of_boundaries", _ip),
    ("stage_edge_values", _dp), ("xmom_edge_values", _dp), ("ymom_edge_values", _dp),
    ("bed_edge_values", _dp), ("height_edge_values", _dp),
    ("stage_centroid_values", _dp), ("xmom_centroid_values", _dp),
    ("ymom_centroid_values", _dp), ("bed_centroid_values", _dp),
    ("height_centroid_values", _dp),
    ("stage_vertex_values", _dp), ("xmom_vertex_values", _dp), ("ymom_vertex_values", _dp),
    ("bed_vertex_values", _dp), ("height_vertex_values", _dp),
    ("stage_boundary_values", _dp), ("xmom_boundary_values", _dp),
    ("ymom_boundary_values", _dp), ("bed_boundary_values", _dp),
    ("stage_explicit_update", _dp), ("xmom_explicit_update", _dp),
    ("ymom_explicit_update", _dp),
    ("flux_update_frequency", _ip), ("update_next_flux", _ip), ("update_extrapolation", _ip),
    ("edge_timestep", _dp), ("edge_flux_work", _dp), ("neigh_work", _dp),
    ("pressuregrad_work", _dp), ("x_centroid_work", _dp), ("y_centroid_work", _dp),
    ("boundary_flux_sum", _dp), ("allow_timestep_increase", _ip),
    ("edge_river_wall_counter", _ip), ("riverwall_elevation", _dp),
    ("riverwall_rowIndex", _ip), ("riverwall_hydraulic_properties", _dp),
    ("stage_semi_implicit_update", _dp), ("xmom_semi_implicit_update", _dp),
    ("ymom_semi_implicit_update", _dp),
]


class _RefDomain(C.Structure):
    _fields_ = _REF_FIELDS


class RefKernels:
    """Kernels of the reference's compiled sw_domain_orig.c and quantity.c.
    `fresh=True` loads a private copy of libref_sw.so, so that its static
    counters (the flux call number, the local-updating step count and
    timestep, sw_domain_orig.c:523, :779-786) start from their initial values."""

    # the shared library's static call counters are process-wide, so is this
    _last_fluxcalls = None

    def __init__(self, fresh=False):
        path = os.path.join(HERE, "_ref", "libref_sw.so")
        if fresh:
            import shutil
            import tempfile
            _load(path)     # the build check and its message
            tmp = tempfile.mkdtemp(prefix="ref_sw_")
            copy = os.path.join(tmp, "libref_sw_%d.so" % id(self))
            shutil.copyfile(path, copy)
            path = copy
            self._last_fluxcalls = None     # this instance's own counters
        sw = _load(path)
        q = _load(os.path.join(HERE, "_ref", "libref_quantity.so"))
        self.sw, self.q = sw, q
        sw._protect_new.restype = C.c_double
        sw._protect_new.argtypes = [C.c_void_p]
        sw._extrapolate_second_order_edge_sw.restype = C.c_int64
        sw._extrapolate_second_order_edge_sw.argtypes = [C.c_void_p]
        sw._compute_fluxes_central.restype = C.c_double
        sw._compute_fluxes_central.argtypes = [C.c_void_p, C.c_double]
        sw._manning_friction_flat.restype = None
        sw._manning_friction_flat.argtypes = [C.c_double, C.c_double, C.c_int64, _dp, _dp,
                                              _dp, _dp, _dp, _dp, _dp]
        sw._manning_friction_sloped.restype = None
        sw._manning_friction_sloped.argtypes = [C.c_double, C.c_double, C.c_int64, _dp, _dp,
                                                _dp, _dp, _dp, _dp, _dp, _dp]
        sw._orig_fix_negative_cells.restype = C.c_int64
        sw._orig_fix_negative_cells.argtypes = [C.c_void_p]
        sw.adjust_edgeflux_with_weir.restype = C.c_double
        sw.adjust_edgeflux_with_weir.argtypes = [_dp] + [C.c_double] * 9 + [_dp]
        sw._flux_function_central.restype = C.c_int64
        sw._flux_function_central.argtypes = [
            _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
            C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp,
            C.c_double, C.c_double, C.c_int64]
        sw._rotate.restype = C.c_int64
        sw._rotate.argtypes = [_dp, C.c_double, C.c_double]
        q._update.restype = C.c_int64
        q._update.argtypes = [C.c_int64, C.c_double, _dp, _dp, _dp]
        q._saxpy_centroid_values.restype = C.c_int64
        q._saxpy_centroid_values.argtypes = [C.c_int64, C.c_double, C.c_double, _dp, _dp]
        q._backup_centroid_values.restype = C.c_int64
        q._backup_centroid_values.argtypes = [C.c_int64, _dp, _dp]
        sw._compute_flux_update_frequency.restype = C.c_int64
        sw._compute_flux_update_frequency.argtypes = [C.c_void_p, C.c_double]
        self.lts_M = 1

    def bind(self, d):
        self.d = d
        N, M = d.N, d.M
        self._buf = dict(
            already_computed_flux=np.zeros(3 * N, np.int64),
            flux_update_frequency=np.ones(3 * N, np.int64),
            update_next_flux=np.ones(3 * N, np.int64),
            update_extrapolation=np.ones(N, np.int64),
            edge_flux_work=np.zeros(9 * N), neigh_work=np.zeros(9 * N),
            pressuregrad_work=np.zeros(3 * N),
            allow_timestep_increase=np.ones(1, np.int64),
            bed_boundary_values=np.zeros(max(M, 1)),
        )
        b = self._buf
        D = _RefDomain()
        D.number_of_elements = N
        D.boundary_length = M
        D.number_of_riverwall_edges = d.number_of_riverwall_edges
        D.ncol_riverwall_hydraulic_properties = d.ncol_riverwall_hydraulic_properties
        D.max_flux_update_frequency = self.lts_M
        D.neighbours = _i(d.neighbours)
        D.neighbour_edges = _i(d.neighbour_edges)
        D.surrogate_neighbours = _i(d.surrogate_neighbours)
        D.normals = _d(d.normals)
        D.edgelengths = _d(d.edgelengths)
        D.radii = _d(d.radii)
        D.areas = _d(d.areas)
        D.edge_flux_type = _i(d.edge_flux_type)
        D.tri_full_flag = _i(d.tri_full_flag)
        D.already_computed_flux = _i(b["already_computed_flux"])
        D.max_speed = _d(d.max_speed)
        D.vertex_coordinates = _d(d.vertex_coordinates)
        D.edge_coordinates = _d(d.edge_midpoints)
        D.centroid_coordinates = _d(d.centroids)
        D.number_of_boundaries = _i(d.number_of_boundaries)
        D.stage_edge_values = _d(d.stage_e)
        D.xmom_edge_values = _d(d.xmom_e)
        D.ymom_edge_values = _d(d.ymom_e)
        D.bed_edge_values = _d(d.elev_e)
        D.height_edge_values = _d(d.height_e)
        D.stage_centroid_values = _d(d.stage_c)
        D.xmom_centroid_values = _d(d.xmom_c)
        D.ymom_centroid_values = _d(d.ymom_c)
        D.bed_centroid_values = _d(d.elev_c)
        D.height_centroid_values = _d(d.height_c)
        D.stage_vertex_values = _d(d.stage_v)
        D.xmom_vertex_values = _d(d.xmom_v)
        D.ymom_vertex_values = _d(d.ymom_v)
        D.bed_vertex_values = _d(d.elev_v)
        D.height_vertex_values = _d(d.height_v)
        D.stage_boundary_values = _d(d.stage_b)
        D.xmom_boundary_values = _d(d.xmom_b)
        D.ymom_boundary_values = _d(d.ymom_b)
        D.bed_boundary_values = _d(b["bed_boundary_values"])
        D.stage_explicit_update = _d(d.stage_ex)
        D.xmom_explicit_update = _d(d.xmom_ex)
        D.ymom_explicit_update = _d(d.ymom_ex)
        D.flux_update_frequency = _i(b["flux_update_frequency"])
        D.update_next_flux = _i(b["update_next_flux"])
        D.update_extrapolation = _i(b["update_extrapolation"])
        D.edge_timestep = _d(d.edge_timestep)
        D.edge_flux_work = _d(b["edge_flux_work"])
        D.neigh_work = _d(b["neigh_work"])
        D.pressuregrad_work = _d(b["pressuregrad_work"])
        D.x_centroid_work = _d(d.x_centroid_work)
        D.y_centroid_work = _d(d.y_centroid_work)
        D.boundary_flux_sum = _d(d.boundary_flux_sum)
        D.allow_timestep_increase = _i(b["allow_timestep_increase"])
        D.edge_river_wall_counter = _i(d.edge_river_wall_counter)
        D.riverwall_elevation = _d(d.riverwall_elevation)
        D.riverwall_rowIndex = _i(d.riverwall_rowIndex)
        D.riverwall_hydraulic_properties = _d(d.riverwall_hydraulic_properties)
        D.stage_semi_implicit_update = _d(d.stage_si)
        D.xmom_semi_implicit_update = _d(d.xmom_si)
        D.ymom_semi_implicit_update = _d(d.ymom_si)
        self.D = D

    def _params(self):
        d, D = self.d, self.D
        D.epsilon = d.epsilon
        D.H0 = d.H0
        D.g = d.g
        D.optimise_dry_cells = 0
        D.evolve_max_timestep = d.evolve_max_timestep
        D.extrapolate_velocity_second_order = d.extrapolate_velocity_second_order
        D.minimum_allowed_height = d.minimum_allowed_height
        D.maximum_allowed_speed = 0.0
        D.low_froude = d.low_froude
        D.beta_w, D.beta_w_dry = d.beta_w, d.beta_w_dry
        D.beta_uh, D.beta_uh_dry = d.beta_uh, d.beta_uh_dry
        D.beta_vh, D.beta_vh_dry = d.beta_vh, d.beta_vh_dry

    def protect(self):
        self._params()
        return self.sw._protect_new(C.byref(self.D))

    def extrapolate(self):
        self._params()
        return self.sw._extrapolate_second_order_edge_sw(C.byref(self.D))

    def compute_fluxes(self, substep, timestep):
        # _compute_fluxes_central derives the substep from static call counters
        # (sw_domain_orig.c:784-807) that reset whenever timestep_fluxcalls
        # changes. Substep 0 is forced by switching between 2 and 3; substep 1
        # keeps the previous value, so (call - base_call) % fluxcalls == 1.
        self._params()
        if substep == 0:
            fc = 3 if self._last_fluxcalls == 2 else 2
        else:
            if self._last_fluxcalls is None:
                raise RuntimeError("substep 1 requested before substep 0")
            fc = self._last_fluxcalls
        self._set_last_fluxcalls(fc)
        self.D.timestep_fluxcalls = fc
        return self.sw._compute_fluxes_central(C.byref(self.D), timestep)

    def _set_last_fluxcalls(self, fc):
        if "_last_fluxcalls" in self.__dict__:
            self._last_fluxcalls = fc
        else:
            RefKernels._last_fluxcalls = fc

    def set_max_flux_update_frequency(self, M):
        """Local flux updating: the reference's arrays as bind() set them
        (their initial values) and max_flux_update_frequency = M. The step
        count is the library's static, so use a fresh instance per run."""
        self.lts_M = M
        self.d.edge_timestep[:] = 1.0e100
        self.bind(self.d)

    def compute_flux_update_frequency(self, timestep):
        self.sw._compute_flux_update_frequency(C.byref(self.D), timestep)

    def friction(self):
        d = self.d
        if d.use_sloped_mannings:
            self.sw._manning_friction_sloped(d.g, d.minimum_allowed_height, d.N,
                                             _d(d.vertex_coordinates), _d(d.stage_c),
                                             _d(d.elev_v), _d(d.xmom_c), _d(d.ymom_c),
                                             _d(d.friction_c), _d(d.xmom_si), _d(d.ymom_si))
            return
        self.sw._manning_friction_flat(d.g, d.minimum_allowed_height, d.N, _d(d.stage_c),
                                       _d(d.elev_c), _d(d.xmom_c), _d(d.ymom_c),
                                       _d(d.friction_c), _d(d.xmom_si), _d(d.ymom_si))

    def update(self, dt):
        d = self.d
        for q, ex, si in ((d.stage_c, d.stage_ex, d.stage_si),
                          (d.xmom_c, d.xmom_ex, d.xmom_si),
                          (d.ymom_c, d.ymom_ex, d.ymom_si)):
            if self.q._update(d.N, dt, _d(q), _d(ex), _d(si)) != 0:
                raise AssertionError("update: division by zero in semi implicit update")

    def fix_negative(self):
        self._params()
        return self.sw._orig_fix_negative_cells(C.byref(self.D))

    def backup(self):
        d = self.d
        for q, b in ((d.stage_c, d.stage_bk), (d.xmom_c, d.xmom_bk), (d.ymom_c, d.ymom_bk)):
            self.q._backup_centroid_values(d.N, _d(q), _d(b))

    def saxpy(self, a, b):
        d = self.d
        for q, bk in ((d.stage_c, d.stage_bk), (d.xmom_c, d.xmom_bk), (d.ymom_c, d.ymom_bk)):
            self.q._saxpy_centroid_values(d.N, a, b, _d(q), _d(bk))

    def flux_function_central(self, normal, ql, qr, h_left, h_right, hle, hre,
                              epsilon, ze, g, low_froude=0):
        ql = np.ascontiguousarray(ql, np.float64)
        qr = np.ascontiguousarray(qr, np.float64)
        ef = np.zeros(3)
        ms = np.zeros(1)
        pf = np.zeros(1)
        self.sw._flux_function_central(_d(ql), _d(qr), h_left, h_right, hle, hre,
                                       float(normal[0]), float(normal[1]), epsilon, ze, 0.0,
                                       g, _d(ef), _d(ms), _d(pf), 0.0, 0.0, low_froude)
        return ef, ms[0], pf[0]

    def rotate(self, q, normal, direction=1):
        r = np.array(q, dtype=np.float64)
        n2 = -normal[1] if direction == -1 else normal[1]
        self.sw._rotate(_d(r), float(normal[0]), float(n2))
        return r

    def adjust_edgeflux_with_weir(self, edgeflux, h_left, h_right, g, weir_height,
                                  Qfactor, s1, s2, h1, h2, max_speed):
        ef = np.array(edgeflux, dtype=np.float64)
        ms = np.array([max_speed], dtype=np.float64)
        self.sw.adjust_edgeflux_with_weir(_d(ef), h_left, h_right, g, weir_height, Qfactor,
                                          s1, s2, h1, h2, _d(ms))
        return ef, ms[0]

"""`Domain`: the anuga.Domain interface of the DE step, running on one MI355X.

This is the host side of the drop-in boundary. It mirrors the parts of
anuga.Domain a DE scenario uses (anuga/shallow_water/shallow_water_domain.py,
anuga/abstract_2d_finite_volumes/generic_domain.py):

  set_flow_algorithm('DE0'|'DE1'), set_quantity, get_quantity,
  set_boundary({tag: boundary}), evolve(yieldstep, finaltime), get_time,
  get_water_volume, timestep bookkeeping (recorded_min/max_timestep,
  number_of_steps), set_fixed_flux_timestep, boundary_flux_integral.

The evolve control flow (yield cadence, update_timestep's CFL / clamping /
small-step logic, generic_domain.py:1715-2051, :2349-2415) runs here on the
host exactly as in the reference, one library call per substep; the state
stays in HBM and only the CFL timestep (8 bytes) crosses per step.
"""

from __future__ import annotations

import ctypes as C
import os
import warnings

import numpy as np

from . import _native
from .boundaries import Boundary
from .mesh import Mesh

# anuga/config.py
G = 9.8
EPSILON = 1.0e-12
MAX_TIMESTEP = 1.0e+3
MIN_TIMESTEP = 1.0e-6
MAX_SMALLSTEPS = 50
H0 = 1.0e-3

_QIDS = {"stage": _native.STAGE, "xmomentum": _native.XMOMENTUM,
         "ymomentum": _native.YMOMENTUM, "elevation": _native.ELEVATION,
         "friction": _native.FRICTION, "height": _native.HEIGHT}

_dp = C.POINTER(C.c_double)


def _dptr(a):
    return a.ctypes.data_as(_dp)


class Quantity:
    """Read view of one quantity of a device-resident Domain.

    centroid_values / edge_values download from the device; vertex_values are
    rebuilt from the edge values as the extrapolation does
    (sw_domain_orig.c:2634-2676, v_i = e_{i+1} + e_{i+2} - e_i).
    """

    def __init__(self, domain, name):
        self.domain = domain
        self.name = name

    @property
    def centroid_values(self):
        return self.domain._get_centroids(self.name)

    @property
    def edge_values(self):
        return self.domain._get_edges(self.name)

    @property
    def vertex_values(self):
        if self.name == "friction":
            c = self.centroid_values
            return np.repeat(c[:, None], 3, axis=1)
        if self.name == "elevation" and not self.domain._distributed:
            return self.domain._elev_vertex.copy()
        e = self.edge_values
        v = np.empty_like(e)
        v[:, 0] = e[:, 1] + e[:, 2] - e[:, 0]
        v[:, 1] = e[:, 0] + e[:, 2] - e[:, 1]
        v[:, 2] = e[:, 0] + e[:, 1] - e[:, 2]
        return v

    def get_integral(self):
        return float(np.sum(self.domain.mesh.areas * self.centroid_values))


MINIMUM_STORABLE_HEIGHT = 1.0e-03    # anuga/config.py minimum_storable_height


def _gauge_row(quantities, w, z, uh, vh):
    """_quantities2csv (anuga/abstract_2d_finite_volumes/gauge.py:30-81)."""
    import math
    out = []
    for q in quantities:
        if q == "stage":
            out.append(w)
        elif q == "elevation":
            out.append(z)
        elif q == "xmomentum":
            out.append(uh)
        elif q == "ymomentum":
            out.append(vh)
        elif q == "depth":
            out.append(w - z)
        elif q == "momentum":
            out.append(math.sqrt(uh ** 2 + vh ** 2))
        elif q == "speed":
            if w - z < 0.001:
                out.append(0.0)
            elif uh < 1.0e6:
                out.append(math.sqrt(uh ** 2 + vh ** 2) / (w - z))
            else:
                out.append(0.0)
        elif q == "bearing":
            out.append(_bearing(uh, vh))
    return out


def _boundary_values(B, t, ids):
    """[stage, xmomentum, ymomentum] of a time-dependent boundary on its
    edges ids at time t: per edge (File/Field boundaries) or one row."""
    if hasattr(B, "values_at"):
        return B.values_at(t, ids)
    return B.values(t)


def _bearing(uh, vh):
    """calc_bearing (anuga/abstract_2d_finite_volumes/util.py:156-171): the
    angle in degrees, in [0, 360), from (0, -1) to the momentum vector
    (numerical_tools.angle); NaN for zero momentum."""
    import math
    if uh == 0 and vh == 0:
        return float("nan")
    n = math.sqrt(uh * uh + vh * vh)
    theta = math.acos(max(-1.0, min(1.0, -vh / n)))
    if uh / n < 0:
        theta = 2 * math.pi - theta
    return math.degrees(theta)


class Domain:
    """Shallow-water domain whose DE step runs on the GPU.

    Domain(coordinates, vertices, boundary) builds the mesh like anuga.Domain;
    `mesh=` accepts a prebuilt anuga_core_amd.mesh.Mesh. `precision` is 'fp64'
    (bit-comparable with the reference) or 'mixed'; `reorder` picks the
    library's internal triangle order: True or 'tiles' (compact tiles from a
    recursive bisection, which let FP64 steps fuse extrapolation and flux), 'rcm' (Reverse
    Cuthill-McKee) or False (the mesh's order). Results are the same bits in
    every order.
    """

    def __init__(self, coordinates=None, vertices=None, boundary=None, mesh=None,
                 use_inscribed_circle=False, precision="fp64", reorder=True,
                 starttime=0.0, verbose=False):
        if mesh is None:
            mesh = Mesh(coordinates, vertices, boundary, use_inscribed_circle)
        self.mesh = mesh
        self.number_of_elements = self.number_of_triangles = N = mesh.number_of_triangles
        self.boundary_length = mesh.boundary_length
        self.verbose = verbose
        self.precision = precision
        # True / "tiles": bisection tiles (the fused FP64 kernel); "rcm":
        # Reverse Cuthill-McKee; False: the mesh's own order
        self._reorder = {True: 2, "tiles": 2, "rcm": 1, False: 0, None: 0}[reorder]

        self.g = G
        self.epsilon = EPSILON
        self.H0 = H0
        self.low_froude = 0
        self.use_sloped_mannings = False
        self.extrapolate_velocity_second_order = 1
        self.evolve_max_timestep = MAX_TIMESTEP
        self.evolve_min_timestep = MIN_TIMESTEP
        self.max_smallsteps = MAX_SMALLSTEPS
        self.fixed_flux_timestep = None
        self.max_flux_update_frequency = 1
        self.centroid_transmissive_bc = False
        self.boundary_map = None
        self.starttime = float(starttime)
        self.relative_time = 0.0
        self.evolved_called = False
        self.smallsteps = 0
        self.timestep = 0.0
        self.flux_timestep = 0.0
        self.number_of_steps = 0
        self.recorded_min_timestep = self.recorded_max_timestep = 0.0
        self.dt_history = []
        self.record_dt_history = True
        # True: update_timestep runs on the device and runs of steps need one
        # host synchronisation each; False: one library call per substep with
        # update_timestep on the host (the reference's sequencing, for checks)
        self.device_timestep = True
        # "fused": each step is a few fused kernels; "per_operation": each step
        # makes the reference's own calls in its order, one library call each
        self.step_mode = "fused"
        self._distributed = False
        self._last_negative_total = 0
        self.mass_error = 0.0
        self._elev_vertex = np.zeros((N, 3))
        # bumped by every adjoint_record, so a gradient can tell which run it has
        self.adjoint_record_serial = 0

        self._handle = C.c_void_p()
        self._lts_on_device = 1
        self._create_device()
        self.set_flow_algorithm("DE0")

        # riverwalls (shallow_water_domain.py:366-379): edge_flux_type 1 on a
        # wall half-edge; the data lives in riverwallData
        from .riverwall import RiverWall
        self.edge_flux_type = np.zeros(3 * N, np.int64)
        self.number_of_riverwall_edges = 0
        self.riverwallData = RiverWall(self)
        self.create_riverwalls = self.riverwallData.create_riverwalls

        # output (shallow_water_domain.py:335-345, anuga/config.py): SWW frames
        # are prepared on the device at yields; storage is off unless set_store
        self.store = False
        self.store_centroids = True
        self.smooth = True
        self.minimum_storable_height = MINIMUM_STORABLE_HEIGHT
        self.quantities_to_be_stored = {"elevation": 1, "friction": 1, "stage": 2,
                                        "xmomentum": 2, "ymomentum": 2}
        self.datadir = "."
        self.simulation_name = "domain"
        self.output_frequency = 1
        self.yieldstep_counter = 0
        self.writer = None
        self._output_nodes = False
        self._gauges = None
        self._time_bc = []
        self._bc_values = None
        self.fractional_step_operators = []

    # -- device handle -----------------------------------------------------------
    def _create_device(self):
        L = _native.lib()
        m = self.mesh
        self._mesh_keep = dict(
            nb=np.ascontiguousarray(m.neighbours, np.int64),
            nbe=np.ascontiguousarray(m.neighbour_edges, np.int64),
            sur=np.ascontiguousarray(m.surrogate_neighbours, np.int64),
            nob=np.ascontiguousarray(m.number_of_boundaries, np.int64),
            nrm=np.ascontiguousarray(m.normals), el=np.ascontiguousarray(m.edgelengths),
            rad=np.ascontiguousarray(m.radii), area=np.ascontiguousarray(m.areas),
            cen=np.ascontiguousarray(m.centroid_coordinates),
            em=np.ascontiguousarray(m.edge_midpoint_coordinates),
            vx=np.ascontiguousarray(m.vertex_coordinates))
        k = self._mesh_keep
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int64))
        desc = _native.AnugaMesh(self.number_of_elements, self.boundary_length,
                                 ip(k["nb"]), ip(k["nbe"]), ip(k["sur"]), ip(k["nob"]),
                                 _dptr(k["nrm"]), _dptr(k["el"]), _dptr(k["rad"]),
                                 _dptr(k["area"]), _dptr(k["cen"]), _dptr(k["em"]),
                                 None if self.precision == "mixed" else _dptr(k["vx"]))
        self._timestep_fluxcalls = 1
        self._params = self._make_params(fp64_defaults=True)
        rc = L.anuga_create(C.byref(desc), C.byref(self._params), C.byref(self._handle))
        if rc != 0:
            msg = L.anuga_last_error(None)
            raise _native.DeviceError("anuga_create failed (%d): %s"
                                      % (rc, msg.decode() if msg else ""))
        del self._mesh_keep

    def _make_params(self, fp64_defaults=False):
        prec = {"fp64": _native.PRECISION_FP64, "mixed": _native.PRECISION_MIXED}[self.precision]
        if fp64_defaults:
            return _native.AnugaParams(G, EPSILON, 1.0e-12, 0.5, 0.0, 0.5, 0.0, 0.5, 0.0,
                                       0, 1, 1, prec, int(self._reorder))
        return _native.AnugaParams(self.g, self.epsilon, self.minimum_allowed_height,
                                   self.beta_w, self.beta_w_dry, self.beta_uh, self.beta_uh_dry,
                                   self.beta_vh, self.beta_vh_dry, self.low_froude,
                                   self.extrapolate_velocity_second_order,
                                   self._timestep_fluxcalls, prec, int(self._reorder),
                                   int(self.use_sloped_mannings))

    def _sync_params(self):
        """The parameters, and max_flux_update_frequency (also when set
        directly, as the reference's C reads it from the domain). Only euler
        steps call compute_flux_update_frequency (generic_domain.py:1960-1963),
        so with rk2 / rk3 the reference's edge state keeps its initial values
        and every flux is evaluated: the library gets 1 then. A change
        restarts the edge state."""
        p = self._make_params()
        M = int(self.max_flux_update_frequency) if self._timestep_fluxcalls == 1 else 1
        if M == 1 and self._lts_on_device != 1:
            self._call("anuga_set_max_flux_update_frequency", 1)
            self._lts_on_device = 1
        self._call("anuga_set_params", C.byref(p))
        if M != self._lts_on_device:
            self._call("anuga_set_max_flux_update_frequency", M)
            self._lts_on_device = M

    def _call(self, name, *args):
        L = _native.lib()
        rc = getattr(L, name)(self._handle, *args)
        _native.check(self._handle, rc, name)

    def close(self):
        """Release the device buffers now (also done when the Domain is
        garbage-collected)."""
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            _native.lib().anuga_destroy(h)
            self._handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return self.number_of_elements

    # -- algorithm presets ---------------------------------------------------------
    # shallow_water_domain.py:583-895: CFL, timestepping, minimum_allowed_height,
    # beta (w, uh, vh), beta_dry
    FLOW_ALGORITHMS = {"DE0": (0.9, "euler", 1.0e-12, 0.5, 0.0),
                       "DE1": (1.0, "rk2", 1.0e-5, 1.0, 0.0),
                       "DE2": (1.0, "rk3", 1.0e-5, 1.0, 0.0),
                       "DE1_7": (1.0, "rk2", 1.0e-12, 0.75, 0.1),
                       "DE0_7": (0.9, "euler", 1.0e-12, 0.7, 0.1)}

    def set_flow_algorithm(self, algorithm="DE0"):
        """DE presets (shallow_water_domain.py:583-895): DE0, DE1, DE2, DE0_7, DE1_7."""
        algorithm = str(algorithm).replace(".", "_")
        if algorithm not in self.FLOW_ALGORITHMS:
            raise Exception("Unknown or unsupported flow_algorithm %s; supported: %s"
                            % (algorithm, ", ".join(self.FLOW_ALGORITHMS)))
        cfl, method, mah, b, bd = self.FLOW_ALGORITHMS[algorithm]
        self.CFL = cfl
        self.set_timestepping_method(method)
        self.minimum_allowed_height = mah
        self.flow_algorithm = algorithm
        self.beta_w = self.beta_uh = self.beta_vh = b
        self.beta_w_dry = self.beta_uh_dry = self.beta_vh_dry = bd
        self.default_order = 2
        self.maximum_allowed_speed = 0.0

    def set_low_froude(self, low_froude=0):
        """Low-Froude momentum diffusion scaling 0, 1 or 2
        (shallow_water_domain.py:1349-1357; FP64 path)."""
        assert low_froude in [0, 1, 2]
        self.low_froude = low_froude

    def set_sloped_mannings_function(self, flag=True):
        """Manning friction over the sloped wetted area
        (shallow_water_domain.py:1159-1169; FP64 path)."""
        self.use_sloped_mannings = bool(flag)

    def get_using_discontinuous_elevation(self):
        return self.flow_algorithm.startswith("DE")

    def _set_riverwalls(self, edge_flux_type, riverwall_elevation, riverwall_rowIndex,
                        hydraulic_properties):
        """Upload riverwall data in the reference's layout (see
        anuga_core_amd.riverwall)."""
        eft = np.ascontiguousarray(edge_flux_type, np.int64).ravel()
        elev = np.ascontiguousarray(riverwall_elevation, np.float64)
        rows = np.ascontiguousarray(riverwall_rowIndex, np.int64)
        hp = np.ascontiguousarray(np.atleast_2d(hydraulic_properties), np.float64)
        if self.precision != "fp64" and len(elev) > 0:
            raise Exception("riverwalls need precision='fp64'")
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int64))
        self._call("anuga_set_riverwalls", ip(eft), len(elev), _dptr(elev), ip(rows),
                   hp.shape[0], hp.shape[1], _dptr(hp))
        self.edge_flux_type = eft.copy()
        self.number_of_riverwall_edges = len(elev)

    def get_flow_algorithm(self):
        return self.flow_algorithm

    def set_local_extrapolation_and_flux_updating(self, nlevels=8):
        """shallow_water_domain.py:1193-1215 with the reference's checks and
        exceptions: each edge's flux is then re-evaluated every 1, 2, 4, ...
        2**nlevels steps by its edge timestep (sw_domain_orig.c:509-669), on
        the device, with host-sequenced euler steps."""
        self.max_flux_update_frequency = 2 ** nlevels
        if self.max_flux_update_frequency != 1:
            if self.timestepping_method != "euler":
                raise Exception("Local extrapolation and flux updating only supported with "
                                "euler timestepping")
            if not self.get_using_discontinuous_elevation():
                raise Exception("Local extrapolation and flux updating only supported for "
                                "discontinuous flow algorithms")
        self._sync_params()

    def compute_flux_update_frequency(self):
        """shallow_water_domain.py:2718-2745 with the step's timestep."""
        self._call("anuga_compute_flux_update_frequency", C.c_double(self.timestep))

    def set_timestepping_method(self, method):
        methods = {"euler": 1, "rk2": 2, "rk3": 3, 1: 1, 2: 2, 3: 3}
        if method not in methods:
            raise Exception("%s is an incorrect timestepping type" % method)
        self._timestep_fluxcalls = methods[method]
        self.timestepping_method = ("euler", "rk2", "rk3")[self._timestep_fluxcalls - 1]

    def get_timestepping_method(self):
        return self.timestepping_method

    def set_CFL(self, cfl):
        self.CFL = float(cfl)

    def get_CFL(self):
        return self.CFL

    def set_minimum_allowed_height(self, h):
        self.minimum_allowed_height = float(h)

    def set_evolve_max_timestep(self, t):
        self.evolve_max_timestep = float(t)

    def set_evolve_min_timestep(self, t):
        self.evolve_min_timestep = float(t)

    def set_fixed_flux_timestep(self, flux_timestep=None):
        """generic_domain.py:2332-2347"""
        if flux_timestep is None:
            self.fixed_flux_timestep = None
        elif flux_timestep > 0.0:
            self.fixed_flux_timestep = float(flux_timestep)
        else:
            raise Exception("flux_timestep needs to be greater than 0.0")

    def set_centroid_transmissive_bc(self, flag):
        self.centroid_transmissive_bc = bool(flag)
        if self.boundary_map is not None:
            self.set_boundary(self.boundary_map)

    # -- output -----------------------------------------------------------------------
    def set_store(self, flag=True):
        """Write an SWW file (datadir/name.sww) at every output yield."""
        self.store = bool(flag)

    def get_store(self):
        return self.store

    def set_store_centroids(self, flag=True):
        self.store_centroids = bool(flag)

    def get_store_centroids(self):
        return self.store_centroids

    def set_store_vertices_uniquely(self, flag=True, reduction=None):
        if flag:
            raise Exception("SWW vertex values are stored smoothly (node averages of the "
                            "centroids); unique vertex storage is not supported")
        self.smooth = True

    def set_store_vertices_smoothly(self, flag=True, reduction=None):
        self.set_store_vertices_uniquely(not flag)

    def set_minimum_storable_height(self, minimum_storable_height):
        self.minimum_storable_height = float(minimum_storable_height)

    def get_minimum_storable_height(self):
        return self.minimum_storable_height

    def set_quantities_to_be_stored(self, q):
        """{name: 1 (static) or 2 (every yield)} (shallow_water_domain.py:1520-1570)."""
        if q is None:
            self.quantities_to_be_stored = {}
            self.store = False
            return
        if isinstance(q, (list, tuple)):
            q = {name: 2 for name in q}
        for name, flag in q.items():
            if name not in ("stage", "xmomentum", "ymomentum", "elevation", "friction"):
                raise Exception("Quantity %s is not a quantity of the domain" % name)
            if flag not in (1, 2):
                raise Exception("Store flag of %s must be 1 (static) or 2 (dynamic)" % name)
        self.quantities_to_be_stored = dict(q)

    def set_datadir(self, name):
        self.datadir = name

    def get_datadir(self):
        return self.datadir

    def set_name(self, name=None, timestamp=False):
        """Base name of the SWW file (generic_domain.py:1565-1596)."""
        if name is None:
            import __main__
            main = getattr(__main__, "__file__", None)
            name = "output_" + os.path.splitext(os.path.basename(main))[0] if main else "domain"
        name = str(name)
        if name.endswith(".sww"):
            name = name[:-4]
        if timestamp:
            import time
            name += time.strftime("_%Y%m%d_%H%M%S")
        self.simulation_name = name

    def get_name(self):
        return self.simulation_name

    def _ensure_output_nodes(self):
        if self._output_nodes:
            return
        from .sww import node_structure
        m = self.mesh
        counts, vvi = node_structure(m.triangles, m.number_of_nodes)
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int64))
        self._call("anuga_set_output_nodes", m.number_of_nodes, ip(counts), ip(vvi))
        self._output_nodes = True

    def _sww_frame(self, minimum_storable_height):
        """(node values [3][nodes], centroid values [3][N]) of stage,
        xmomentum, ymomentum as float32, prepared on the device."""
        self._ensure_output_nodes()
        nv = np.empty((3, self.mesh.number_of_nodes), np.float32)
        cv = np.empty((3, self.number_of_elements), np.float32)
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        self._call("anuga_get_sww_frame", C.c_double(minimum_storable_height), fp(nv), fp(cv))
        return nv, cv

    def _static_output(self, name):
        self._ensure_output_nodes()
        nv = np.empty(self.mesh.number_of_nodes, np.float32)
        cv = np.empty(self.number_of_elements, np.float32)
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        self._call("anuga_get_static_output", _QIDS[name], fp(nv), fp(cv))
        return nv, cv

    def initialise_storage(self):
        """Create the SWW writer and store the mesh and static quantities."""
        from .sww import SWW_file
        if self.writer is not None:
            self.writer.close()
        self.writer = SWW_file(self)
        self.writer.store_connectivity()

    def store_timestep(self):
        self.writer.store_timestep()

    # -- gauges -------------------------------------------------------------------------
    def set_gauges(self, points=None, triangles=None, names=None, capacity=4096):
        """Record the centroid state of the triangles holding `points` (or of
        `triangles`) after every step, on the device."""
        if (points is None) == (triangles is None):
            raise Exception("give either points or triangles")
        if points is not None:
            pts = np.atleast_2d(np.asarray(points, np.float64))
            triangles = self.mesh.find_triangles(pts)
            if np.any(triangles < 0):
                raise Exception("gauge point(s) %s fall outside the mesh"
                                % pts[triangles < 0].tolist())
        tri = np.ascontiguousarray(triangles, np.int64).ravel()
        if names is None:
            names = ["gauge%d" % i for i in range(len(tri))]
        self._call("anuga_set_gauges", len(tri), tri.ctypes.data_as(C.POINTER(C.c_int64)),
                   int(capacity))
        self._gauges = None if len(tri) == 0 else dict(
            triangles=tri, names=list(names), capacity=int(capacity), times=[], values=[],
            pending=0)

    def _drain_gauges(self):
        g = self._gauges
        if g is None:
            return
        cap, G = g["capacity"], len(g["triangles"])
        t = np.empty(cap)
        v = np.empty((cap, G, 4))
        n = C.c_int64()
        self._call("anuga_get_gauge_samples", cap, C.byref(n), _dptr(t), _dptr(v))
        g["pending"] = 0
        if n.value:
            g["times"].append(t[:n.value] + self.starttime)
            g["values"].append(v[:n.value].copy())

    def get_gauge_timeseries(self):
        """{'time', 'stage', 'xmomentum', 'ymomentum', 'elevation', 'depth'}:
        arrays (samples, gauges) of every step since set_gauges."""
        g = self._gauges
        if g is None:
            raise Exception("no gauges set")
        self._drain_gauges()
        G = len(g["triangles"])
        t = np.concatenate(g["times"]) if g["times"] else np.empty(0)
        v = np.concatenate(g["values"]) if g["values"] else np.empty((0, G, 4))
        g["times"], g["values"] = [t], [v]
        return dict(time=t, stage=v[:, :, 0], xmomentum=v[:, :, 1], ymomentum=v[:, :, 2],
                    elevation=v[:, :, 3], depth=v[:, :, 0] - v[:, :, 3])

    def write_gauge_csv(self, directory=".", out_name="gauge_",
                        quantities=("stage", "depth", "elevation", "xmomentum", "ymomentum")):
        """One CSV per gauge in the column layout of sww2csv_gauges
        (anuga/abstract_2d_finite_volumes/gauge.py:30-266): time, hours, then
        the quantities (stage, elevation, xmomentum, ymomentum, depth,
        momentum, speed, bearing)."""
        import csv
        ts = self.get_gauge_timeseries()
        paths = []
        for i, name in enumerate(self._gauges["names"]):
            path = os.path.join(directory, out_name + name + ".csv")
            with open(path, "w", newline="") as f:
                w = csv.writer(f)
                w.writerow(["time", "hours"] + list(quantities))
                for r, t in enumerate(ts["time"]):
                    w.writerow([t, t / 3600.0] + _gauge_row(quantities, ts["stage"][r, i],
                                                             ts["elevation"][r, i],
                                                             ts["xmomentum"][r, i],
                                                             ts["ymomentum"][r, i]))
            paths.append(path)
        return paths

    # -- quantities -----------------------------------------------------------------
    def set_quantity(self, name, numeric=None, function=None, location="vertices",
                     indices=None):
        """Quantity.set_values for constants, arrays and f(x, y) (quantity.py:709-1175).

        location='vertices' (the reference default) sets vertex values and
        derives centroids as (v0 + v1 + v2) / 3.0 (quantity.c:665-688);
        location='centroids' sets centroid values directly.
        """
        if name not in ("stage", "xmomentum", "ymomentum", "elevation", "friction"):
            raise Exception("Unknown quantity %s" % name)
        if indices is not None:
            raise Exception("indices are not supported")
        value = function if function is not None else numeric
        if value is None:
            raise Exception("one of numeric or function must be given")
        N = self.number_of_elements
        if location == "centroids":
            if callable(value):
                cen = self.mesh.centroid_coordinates
                c = np.asarray(value(cen[:, 0], cen[:, 1]), np.float64)
                c = np.broadcast_to(c, (N,)).copy()
            else:
                c = np.broadcast_to(np.asarray(value, np.float64), (N,)).copy()
            vtx = np.repeat(c[:, None], 3, axis=1)
        elif location == "vertices":
            if callable(value):
                V = self.mesh.vertex_coordinates
                v = np.asarray(value(V[:, 0], V[:, 1]), np.float64)
                if v.ndim == 0:
                    vtx = np.full((N, 3), float(v))
                else:
                    vtx = v.reshape(N, 3)
            else:
                a = np.asarray(value, np.float64)
                if a.ndim == 0:
                    vtx = np.full((N, 3), float(a))
                elif a.shape == (N, 3):
                    vtx = a.copy()
                else:
                    raise Exception("vertex values must be a constant or an (N,3) array")
            c = (vtx[:, 0] + vtx[:, 1] + vtx[:, 2]) / 3.0
        else:
            raise Exception("Invalid location: %s" % location)
        if name == "elevation":
            self._elev_vertex = np.ascontiguousarray(vtx)
        c = np.ascontiguousarray(c, np.float64)
        self._call("anuga_set_centroid_values", _QIDS[name], _dptr(c))
        self._distributed = False

    def get_quantity(self, name):
        if name not in _QIDS:
            raise KeyError(name)
        return Quantity(self, name)

    @property
    def quantities(self):
        return {n: Quantity(self, n) for n in _QIDS}

    def _get_centroids(self, name):
        out = np.empty(self.number_of_elements)
        self._call("anuga_get_centroid_values", _QIDS[name], _dptr(out))
        return out

    def _get_edges(self, name):
        if name == "friction":
            raise Exception("friction has no edge values")
        out = np.empty((self.number_of_elements, 3))
        self._call("anuga_get_edge_values", _QIDS[name], _dptr(out))
        return out

    def get_centroid_coordinates(self):
        return self.mesh.centroid_coordinates

    def get_vertex_coordinates(self):
        return self.mesh.vertex_coordinates

    # -- boundaries -------------------------------------------------------------------
    def get_boundary_tags(self):
        return self.mesh.get_boundary_tags()

    def set_boundary(self, boundary_map):
        """Bind boundary objects to tags (generic_domain.py set_boundary). Tags
        mapped to None keep their previous kind (the reference skips them)."""
        tags = self.mesh.get_boundary_tags()
        for tag in boundary_map:
            if tag not in tags:
                raise Exception("Tag %s provided does not exist in the domain. "
                                "All boundary tags are %s" % (tag, tags))
        for tag in tags:
            if tag not in boundary_map:
                raise Exception("Tag %s has not been bound to a boundary object" % tag)
        M = self.boundary_length
        kind = np.zeros(max(M, 1), np.int32)
        vals = np.zeros((max(M, 1), 3))
        time_bc = []
        for tag, B in boundary_map.items():
            if B is None:
                continue
            if not isinstance(B, Boundary):
                raise Exception("unsupported boundary object %r" % (B,))
            ids = np.asarray(self.mesh.tag_boundary_cells[tag], np.int64)
            code = B.code
            if code == 1 and self.centroid_transmissive_bc:
                code = 3
            kind[ids] = code
            if getattr(B, "time_dependent", False):
                time_bc.append((B, ids))
                vals[ids] = _boundary_values(B, self.get_time(), ids)
            elif code == 2:
                vals[ids] = B.dirichlet_values[:3]
        self.boundary_map = dict(boundary_map)
        self._bc_values = vals
        self._time_bc = time_bc
        self._call("anuga_set_boundary", kind.ctypes.data_as(C.POINTER(C.c_int32)), _dptr(vals))

    def _update_boundary_slot(self, slot, t):
        """Time-dependent boundary values at time t for one substep slot
        (0: the step's start, 1: t + dt, 2: t + dt/2)."""
        if not self._time_bc:
            return
        vals = self._bc_values
        for B, ids in self._time_bc:
            vals[ids] = _boundary_values(B, t, ids)
        self._call("anuga_set_boundary_values", slot, _dptr(vals))

    # -- step -------------------------------------------------------------------------
    def distribute_to_vertices_and_edges(self):
        """protect + extrapolation on the device, written back to the centroids."""
        self._sync_params()
        me = C.c_double()
        self._call("anuga_distribute_to_vertices_and_edges", C.byref(me))
        self.mass_error += me.value
        if me.value > 0.0 and self.verbose:
            print("Cumulative mass protection: {0} m^3".format(me.value))
        self._distributed = True

    def update_boundary(self):
        """Boundary states are evaluated inside the flux kernel from the edge
        values. In per-operation steps the time-dependent boundary values of
        the current time go to the slot that compute_fluxes reads
        (generic_domain.py:2288-2306)."""
        if self.step_mode == "per_operation":
            self._update_boundary_slot(0, self.get_time())
            self._call("anuga_update_boundary")

    # -- the reference's per-operation calls --------------------------------------------
    def set_step_mode(self, mode="fused"):
        """'fused' (default): a step is a few fused kernels, and with nothing
        time-dependent on the host, runs of steps keep update_timestep on the
        device. 'per_operation': every step makes the calls of
        evolve_one_{euler,rk2,rk3}_step (generic_domain.py:1914-2179) in the
        reference's order, one library call each (backup, distribute,
        update_boundary, compute_fluxes, compute_forcing_terms,
        update_timestep, update_conserved_quantities, saxpy). Both give the
        same bits."""
        if mode not in ("fused", "per_operation"):
            raise ValueError("step mode must be 'fused' or 'per_operation'")
        self.step_mode = mode

    def backup_conserved_quantities(self):
        """generic_domain.py:2187-2191"""
        self._call("anuga_backup_conserved_quantities")

    def saxpy_conserved_quantities(self, a, b):
        """generic_domain.py:2193-2199"""
        self._call("anuga_saxpy_conserved_quantities", C.c_double(a), C.c_double(b))

    def compute_fluxes(self):
        """shallow_water_domain.py:1828-1879: compute_fluxes_ext_central over
        the edge values of the last distribute_to_vertices_and_edges."""
        ft = C.c_double()
        self._call("anuga_compute_fluxes_ext_central", C.c_double(self.evolve_max_timestep),
                   C.byref(ft))
        self.flux_timestep = ft.value

    def compute_forcing_terms(self):
        """generic_domain.py:2417-2427 with the DE forcing term, Manning
        friction (friction.py:22-76)."""
        self._call("anuga_compute_forcing_terms")

    def update_conserved_quantities(self):
        """shallow_water_domain.py:2089-2163: Stage/Xmom/Ymom.update, then
        fix_negative_cells with the reference's warning."""
        self._call("anuga_update_centroids", C.c_double(self.timestep))
        n = C.c_int64()
        self._call("anuga_fix_negative_cells", C.byref(n))
        if n.value > 0:
            warnings.warn("Negative cells being set to zero depth, possible loss of "
                          "conservation. \nConsider using "
                          "domain.report_water_volume_statistics() to check the extent "
                          "of the problem")

    def _per_operation_euler(self, yieldstep, finaltime):
        """generic_domain.py:1914-1972"""
        self.distribute_to_vertices_and_edges()
        self.update_boundary()
        self.compute_fluxes()
        self.compute_forcing_terms()
        self.update_timestep(yieldstep, finaltime)
        if self.max_flux_update_frequency != 1:
            self.compute_flux_update_frequency()
        self.update_conserved_quantities()

    def _per_operation_rk2(self, yieldstep, finaltime):
        """generic_domain.py:1974-2051"""
        self.backup_conserved_quantities()
        self.distribute_to_vertices_and_edges()
        self.update_boundary()
        self.compute_fluxes()
        self.compute_forcing_terms()
        self.update_timestep(yieldstep, finaltime)
        self.update_conserved_quantities()
        self.set_relative_time(self.get_relative_time() + self.timestep)
        self.distribute_to_vertices_and_edges()
        self.update_boundary()
        self.compute_fluxes()
        self.compute_forcing_terms()
        self.update_conserved_quantities()
        self.saxpy_conserved_quantities(0.5, 0.5)

    def _per_operation_rk3(self, yieldstep, finaltime):
        """generic_domain.py:2053-2179"""
        self.backup_conserved_quantities()
        initial_time = self.get_relative_time()
        self.distribute_to_vertices_and_edges()
        self.update_boundary()
        self.compute_fluxes()
        self.compute_forcing_terms()
        self.update_timestep(yieldstep, finaltime)
        self.update_conserved_quantities()
        self.set_relative_time(self.relative_time + self.timestep)
        self.distribute_to_vertices_and_edges()
        self.update_boundary()
        self.compute_fluxes()
        self.compute_forcing_terms()
        self.update_conserved_quantities()
        self.saxpy_conserved_quantities(0.25, 0.75)
        self.set_relative_time(initial_time + self.timestep * 0.5)
        self.distribute_to_vertices_and_edges()
        self.update_boundary()
        self.compute_fluxes()
        self.compute_forcing_terms()
        self.update_conserved_quantities()
        self.saxpy_conserved_quantities(2.0, 1.0)
        self._call("anuga_divide_conserved_quantities", C.c_double(3.0))
        self.set_relative_time(initial_time + self.timestep)

    def update_timestep(self, yieldstep, finaltime):
        """generic_domain.py:2349-2415"""
        if self.fixed_flux_timestep is not None:
            self.flux_timestep = self.fixed_flux_timestep
        timestep = min(self.CFL * self.flux_timestep, self.evolve_max_timestep)
        self.recorded_max_timestep = max(timestep, self.recorded_max_timestep)
        self.recorded_min_timestep = min(timestep, self.recorded_min_timestep)
        if timestep < self.evolve_min_timestep:
            self.smallsteps += 1
            if self.smallsteps > self.max_smallsteps:
                self.smallsteps = 0
                if self._order_ == 1:
                    raise Exception("WARNING: Too small timestep %.16f reached even after %d "
                                    "steps of 1 order scheme" % (timestep, self.max_smallsteps))
                self._order_ = 1
        else:
            self.smallsteps = 0
            if self._order_ == 1 and self.default_order == 2:
                self._order_ = 2
        if self.relative_finaltime is not None and \
                self.relative_time + timestep > self.relative_finaltime:
            timestep = self.relative_finaltime - self.relative_time
        if self.relative_time + timestep > self.relative_yieldtime:
            timestep = self.relative_yieldtime - self.relative_time
        self.timestep = timestep

    def _compute_fluxes(self):
        ft = C.c_double()
        self._call("anuga_compute_fluxes", C.byref(ft))
        self.flux_timestep = ft.value
        self._check_negative_cells()

    def _check_negative_cells(self):
        tot = C.c_int64()
        self._call("anuga_get_negative_cells", C.byref(tot))
        if tot.value > self._last_negative_total:
            self._last_negative_total = tot.value
            warnings.warn("Negative cells being set to zero depth, possible loss of "
                          "conservation. \nConsider using "
                          "domain.report_water_volume_statistics() to check the extent "
                          "of the problem")

    def evolve_one_euler_step(self, yieldstep, finaltime):
        """generic_domain.py:1914-1972"""
        if self.step_mode == "per_operation":
            return self._per_operation_euler(yieldstep, finaltime)
        self._update_boundary_slot(0, self.get_time())
        self._compute_fluxes()
        self.update_timestep(yieldstep, finaltime)
        if self.max_flux_update_frequency != 1:
            self.compute_flux_update_frequency()
        self._call("anuga_update_conserved_quantities", C.c_double(self.timestep), None)

    def evolve_one_rk2_step(self, yieldstep, finaltime):
        """generic_domain.py:1974-2051; backup is the untouched step state."""
        if self.step_mode == "per_operation":
            return self._per_operation_rk2(yieldstep, finaltime)
        self._update_boundary_slot(0, self.get_time())
        self._compute_fluxes()
        self.update_timestep(yieldstep, finaltime)
        self._call("anuga_update_conserved_quantities", C.c_double(self.timestep), None)
        self.relative_time = self.relative_time + self.timestep
        self._update_boundary_slot(1, self.get_time())
        self._call("anuga_rk2_second_substep", C.c_double(self.timestep), None)

    def evolve_one_rk3_step(self, yieldstep, finaltime):
        """generic_domain.py:2053-2179; the backup is the untouched step state."""
        if self.step_mode == "per_operation":
            return self._per_operation_rk3(yieldstep, finaltime)
        initial = self.relative_time
        self._update_boundary_slot(0, self.get_time())
        self._compute_fluxes()
        self.update_timestep(yieldstep, finaltime)
        self._call("anuga_update_conserved_quantities", C.c_double(self.timestep), None)
        self._update_boundary_slot(1, self.starttime + (initial + self.timestep))
        self._update_boundary_slot(2, self.starttime + (initial + self.timestep * 0.5))
        self._call("anuga_rk3_later_substeps", C.c_double(self.timestep), None)
        self.relative_time = initial + self.timestep

    def get_time(self):
        return self.starttime + self.relative_time

    def get_relative_time(self):
        return self.relative_time

    def set_relative_time(self, time=0.0):
        """generic_domain.py set_relative_time: time since starttime."""
        self.relative_time = float(time)

    def set_time(self, time=0.0):
        self.relative_time = float(time) - self.starttime

    def get_starttime(self):
        return self.starttime

    def set_starttime(self, time):
        """generic_domain.py:1607-1615"""
        if self.evolved_called:
            raise Exception("Can't change simulation start time once evolve has been called")
        self.starttime = float(time)
        # starttime is now the origin of relative_time
        self.relative_time = 0.0

    def evolve(self, yieldstep=None, finaltime=None, duration=None, skip_initial_step=False,
               outputstep=None):
        """Generator over yield times (shallow_water_domain.py:2300-2407 and
        generic_domain.py:1715-1912; SWW storage at every outputstep, no
        checkpointing)."""
        if self.boundary_map is None:
            raise AssertionError("Boundary tags must be bound to boundary objects before "
                                 "evolving system, e.g. using the method set_boundary.\n"
                                 "This system has the boundary tags %s "
                                 % self.get_boundary_tags())
        if not (0 <= self.beta_w <= 2.0):
            raise AssertionError("Attribute self.beta_w must be in the interval [0, 2]")
        if outputstep is None:
            outputstep = yieldstep
        if yieldstep is None:
            self.output_frequency = 1
        else:
            freq = outputstep / yieldstep
            if not float(freq).is_integer():
                raise AssertionError("outputstep (%s) should be an integer multiple of "
                                     "yieldstep (%s)" % (outputstep, yieldstep))
            self.output_frequency = int(freq)
        self._sync_params()
        self.distribute_to_vertices_and_edges()
        if self.store and (self.get_relative_time() == 0.0 or not self.evolved_called):
            self.initialise_storage()
        if self.evolved_called:
            skip_initial_step = True
        self.evolved_called = True
        yieldstep = self.evolve_max_timestep if yieldstep is None else float(yieldstep)
        self.yieldstep = yieldstep
        self._order_ = self.default_order
        if finaltime is not None and duration is not None:
            raise Exception("Only one of finaltime and duration may be specified")
        self.relative_finaltime = None
        if finaltime is not None:
            self.finaltime = float(finaltime)
            self.relative_finaltime = self.finaltime - self.starttime
        if duration is not None:
            self.finaltime = float(duration) + self.get_time()
            self.relative_finaltime = float(duration) + self.relative_time
        if self.relative_finaltime is not None and self.relative_finaltime < self.relative_time:
            warnings.warn("\n finaltime %g is less than current time %g! finaltime set to "
                          "current time" % (self.finaltime, self.get_time()))
            self.relative_finaltime = self.relative_time
            return
        self.relative_yieldtime = self.relative_time + yieldstep
        self.recorded_min_timestep = self.evolve_max_timestep
        self.recorded_max_timestep = self.evolve_min_timestep
        self.number_of_steps = 0
        if not skip_initial_step:
            self.distribute_to_vertices_and_edges()
            yield self._yield()
        while True:
            if self._device_stepping():
                reached = self._device_steps(self._chunk_estimate())
            else:
                reached = self._host_step(yieldstep)
            if reached == 1:
                self.relative_time = self.relative_finaltime
                self.distribute_to_vertices_and_edges()
                self._check_negative_cells()
                yield self._yield()
                break
            if reached == 2:
                self.distribute_to_vertices_and_edges()
                self._check_negative_cells()
                yield self._yield()
                self.relative_yieldtime += yieldstep
                self.recorded_min_timestep = self.evolve_max_timestep
                self.recorded_max_timestep = self.evolve_min_timestep
                self.number_of_steps = 0

    def _device_stepping(self):
        """Runs of steps on the device unless something needs the host every
        step (time-dependent boundary values or operator rates)."""
        return (self.device_timestep and self.step_mode == "fused" and not self._time_bc and
                (self.max_flux_update_frequency == 1 or self._timestep_fluxcalls != 1) and
                not any(op.time_dependent for op in self.fractional_step_operators))

    def set_fractional_step_operator(self, operator):
        if operator not in self.fractional_step_operators:
            self.fractional_step_operators.append(operator)

    def get_fractional_step_volume_integral(self):
        """Water added by the fractional-step operators since creation."""
        return float(sum(op.get_cumulative_influx() for op in self.fractional_step_operators))

    def _yield(self):
        """Store the frame of this yield (every output_frequency-th) and
        return the time to yield."""
        if self.store and self.yieldstep_counter % self.output_frequency == 0:
            self.store_timestep()
        self.yieldstep_counter += 1
        return self.get_time()

    def _host_step(self, yieldstep):
        """One step with update_timestep on the host; returns 1 / 2 / 0 as the
        step ends at the final time / a yield time / neither."""
        initial = self.relative_time
        self._distributed = False
        if self._timestep_fluxcalls == 1:
            self.evolve_one_euler_step(yieldstep, self.relative_finaltime)
        elif self._timestep_fluxcalls == 2:
            self.evolve_one_rk2_step(yieldstep, self.relative_finaltime)
        else:
            self.evolve_one_rk3_step(yieldstep, self.relative_finaltime)
        if self.step_mode == "per_operation":
            # the boundary flux integral operator, the first fractional step
            # (shallow_water_domain.py:386-387; the fused steps integrate in
            # their update kernels)
            self._call("anuga_boundary_flux_integral_operator", C.c_double(self.timestep))
        if self.fractional_step_operators:
            # apply_fractional_steps at the step's time as the reference has
            # it then (start for euler, start + dt for rk2 and rk3), in order
            t = self.get_time()
            for op in self.fractional_step_operators:
                if getattr(op, "host_side", False):
                    op.apply(t, self.timestep)
                    continue
                if op.time_dependent:
                    op.update(t, self.timestep)
                self._call("anuga_apply_operator", op._id.value)
        self.relative_time = initial + self.timestep
        if self.record_dt_history:
            self.dt_history.append(self.timestep)
        self.number_of_steps += 1
        if self._gauges is not None:
            g = self._gauges
            self._call("anuga_sample_gauges", C.c_double(self.relative_time))
            g["pending"] += 1
            if g["pending"] == g["capacity"]:
                self._drain_gauges()
        if self.relative_finaltime is not None and \
                self.relative_time >= self.relative_finaltime - EPSILON:
            if self.relative_time > self.relative_finaltime:
                raise Exception("WARNING (domain.py): time overshot finaltime. ")
            return 1
        if self.relative_time >= self.relative_yieldtime:
            return 2
        return 0

    def _chunk_estimate(self):
        """Steps to the next yield or final time from the last timestep, so that
        few launches of a run become no-ops."""
        target = self.relative_yieldtime
        if self.relative_finaltime is not None:
            target = min(target, self.relative_finaltime)
        if self.timestep > 0.0:
            est = (target - self.relative_time) / self.timestep
            return int(max(1, min(256, est + 2)))
        return 8

    def _device_steps(self, nsteps):
        """Up to nsteps steps with update_timestep on the device; returns the
        `reached` code of anuga_evolve_steps."""
        inf = float("inf")
        tc = _native.TimeControl(self.CFL, self.evolve_max_timestep, self.evolve_min_timestep,
                                 inf if self.relative_finaltime is None else self.relative_finaltime,
                                 self.relative_yieldtime,
                                 -1.0 if self.fixed_flux_timestep is None
                                 else self.fixed_flux_timestep,
                                 EPSILON, self.max_smallsteps, self.default_order)
        st = _native.StepState(self.relative_time, self.timestep, self.flux_timestep,
                               self.recorded_min_timestep, self.recorded_max_timestep,
                               self.smallsteps, self._order_, 0, 0)
        hist = np.empty(max(int(nsteps), 1)) if self.record_dt_history else None
        self._distributed = False
        self._call("anuga_evolve_steps", int(nsteps), C.byref(tc), C.byref(st),
                   None if hist is None else _dptr(hist))
        n = st.steps_taken
        self.relative_time = st.relative_time
        self.timestep = st.timestep
        self.flux_timestep = st.flux_timestep
        self.recorded_min_timestep = st.recorded_min_timestep
        self.recorded_max_timestep = st.recorded_max_timestep
        self.smallsteps = st.smallsteps
        self._order_ = st.order
        self.number_of_steps += n
        if hist is not None:
            self.dt_history.extend(hist[:n].tolist())
        self._check_negative_cells()
        self._drain_gauges()
        return st.reached

    def evolve_steps(self, nsteps, finaltime=None):
        """Run exactly `nsteps` steps with no yields (dt still follows
        update_timestep; only finaltime clamps it)."""
        self._sync_params()
        if not self.evolved_called:
            self.distribute_to_vertices_and_edges()
            self.evolved_called = True
        self._order_ = self.default_order
        self.relative_finaltime = None if finaltime is None else float(finaltime) - self.starttime
        self.relative_yieldtime = float("inf")
        self.recorded_min_timestep = self.evolve_max_timestep
        self.recorded_max_timestep = self.evolve_min_timestep
        done = 0
        while done < nsteps:
            if self._device_stepping():
                before = self.number_of_steps
                reached = self._device_steps(nsteps - done)
                done += self.number_of_steps - before
            else:
                reached = self._host_step(float("inf"))
                done += 1
            if reached:
                break

    # -- diagnostics --------------------------------------------------------------------
    def get_water_volume(self):
        v = C.c_double()
        self._call("anuga_get_water_volume", C.byref(v))
        return v.value

    @property
    def boundary_flux_integral(self):
        v = C.c_double()
        self._call("anuga_get_boundary_flux_integral", C.byref(v))
        return v.value

    def get_boundary_flux_integral(self):
        return self.boundary_flux_integral

    def synchronize(self):
        self._call("anuga_synchronize")

    # -- gradients ---------------------------------------------------------------------
    def adjoint_record(self, nsteps, dt, gauges, obs_every=1, checkpoint_every=100):
        """Run `nsteps` DE1 steps of fixed timestep `dt` from the current
        centroid values (set_fixed_flux_timestep semantics with CFL 1; no
        initial distribute), keeping a checkpoint every `checkpoint_every`
        steps on the device, and return the stage of the `gauges` triangles
        after every `obs_every`-th step, shape (nsteps // obs_every, len(gauges)).
        The domain's state advances to the end of the run."""
        self._sync_params()
        g = np.ascontiguousarray(gauges, np.int64)
        nobs = int(nsteps) // int(obs_every)
        out = np.zeros((max(nobs, 1), len(g)))
        self._call("anuga_adjoint_record", int(nsteps), C.c_double(dt), int(checkpoint_every),
                   len(g), g.ctypes.data_as(C.POINTER(C.c_int64)), int(obs_every), _dptr(out))
        self.relative_time += int(nsteps) * float(dt)
        self._adjoint_shape = (nobs, len(g))
        self.adjoint_record_serial += 1
        return out[:nobs]

    def adjoint_gradient(self, dloss_dstage):
        """Discrete adjoint of the last adjoint_record run: given dL/d(stage at
        the gauges) in the record's shape, return (dL/dfriction, dL/delevation,
        dL/d(initial stage, xmomentum, ymomentum)) per triangle; the bed
        derivative is at fixed stage."""
        N = self.number_of_elements
        dl = np.ascontiguousarray(dloss_dstage, np.float64).reshape(self._adjoint_shape)
        gn, gz, gs = np.zeros(N), np.zeros(N), np.zeros((3, N))
        self._call("anuga_adjoint_gradient", _dptr(dl), _dptr(gn), _dptr(gz), _dptr(gs))
        return gn, gz, gs

    def kernel_timing(self, name):
        ms = C.c_double()
        n = C.c_int64()
        self._call("anuga_kernel_timing", name.encode(), C.byref(ms), C.byref(n))
        return ms.value, n.value

    def reset_kernel_timing(self, enable=True):
        self._call("anuga_reset_kernel_timing", int(enable))

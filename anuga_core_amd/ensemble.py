"""`Ensemble`: many independent DE scenarios stepped together on one GPU.

A scenario is what one anuga.Domain holds for a DE0/DE1 run: a mesh, the
centroid quantities (stage, elevation, friction, momenta), boundary
conditions per tag, and its own time and timestep control
(generic_domain.py:1715-2051, update_timestep :2349-2415). With the reference
an ensemble (flood scenarios, a friction sweep) is a loop over Domain objects.
Here all scenarios live in one device handle: every kernel launch and every
replay of the captured step advances all of them, each with its own CFL
timestep, so that thousands of small scenarios fill the GPU together. Each
scenario's results are bit-identical to a Domain of the same inputs.

    ens = Ensemble([mesh_a, mesh_b, ...], flow_algorithm='DE1')
    ens.set_scenario(0, mesh_a, {'elevation': z, 'stage': w, 'friction': 0.03},
                     {'left': 'transmissive', 'right': 'reflective', ...})
    ...
    ens.evolve_steps(500)                 # 500 steps of every scenario
    w0 = ens.get_quantity(0, 'stage')     # centroid values, caller's order
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native
from .boundaries import Boundary, KIND_DIRICHLET, KIND_REFLECTIVE, KIND_TRANSMISSIVE, \
    KIND_TRANSMISSIVE_CENTROID
from .domain import EPSILON, G, MAX_SMALLSTEPS, MAX_TIMESTEP, MIN_TIMESTEP, _QIDS, _dptr

_KINDS = {"reflective": KIND_REFLECTIVE, "transmissive": KIND_TRANSMISSIVE,
          "transmissive_centroid": KIND_TRANSMISSIVE_CENTROID}

# DE presets (shallow_water_domain.py:583-895): CFL, fluxcalls, mah, beta, beta_dry
_PRESETS = {"DE0": (0.9, 1, 1.0e-12, 0.5, 0.0), "DE1": (1.0, 2, 1.0e-5, 1.0, 0.0),
            "DE2": (1.0, 3, 1.0e-5, 1.0, 0.0), "DE1_7": (1.0, 2, 1.0e-12, 0.75, 0.1),
            "DE0_7": (0.9, 1, 1.0e-12, 0.7, 0.1)}


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


class Ensemble:
    """S independent scenarios of one flow algorithm and precision.

    `meshes` gives each scenario's size: anuga_core_amd.mesh.Mesh objects or
    (number_of_triangles, boundary_length) pairs. Each scenario is then filled
    with set_scenario (mesh, quantities, boundaries), in any order.
    """

    def __init__(self, meshes, flow_algorithm="DE1", precision="fp64", reorder=True):
        sizes = []
        for m in meshes:
            if hasattr(m, "number_of_triangles"):
                sizes.append((m.number_of_triangles, m.boundary_length))
            else:
                sizes.append((int(m[0]), int(m[1])))
        if not sizes:
            raise ValueError("an ensemble needs at least one scenario")
        self.size = S = len(sizes)
        self.number_of_triangles = np.array([s[0] for s in sizes], np.int64)
        self.boundary_length = np.array([s[1] for s in sizes], np.int64)
        if flow_algorithm not in _PRESETS:
            raise Exception("Unknown or unsupported flow_algorithm %s; supported: %s"
                            % (flow_algorithm, ", ".join(_PRESETS)))
        self.flow_algorithm = flow_algorithm
        self.CFL, self._fluxcalls, self.minimum_allowed_height, beta, bdry = \
            _PRESETS[flow_algorithm]
        self.precision = precision
        prec = {"fp64": _native.PRECISION_FP64, "mixed": _native.PRECISION_MIXED}[precision]
        self._params = _native.AnugaParams(G, EPSILON, self.minimum_allowed_height,
                                           beta, bdry, beta, bdry, beta, bdry, 0, 1,
                                           self._fluxcalls, prec,
                                           {True: 2, "tiles": 2, "rcm": 1, False: 0, None: 0}[reorder])
        self.evolve_max_timestep = MAX_TIMESTEP
        self.evolve_min_timestep = MIN_TIMESTEP
        self.max_smallsteps = MAX_SMALLSTEPS
        self.default_order = 2
        self._handle = C.c_void_p()
        L = _native.lib()
        rc = L.anuga_ensemble_create(S, _ip(self.number_of_triangles),
                                     _ip(self.boundary_length), C.byref(self._params),
                                     C.byref(self._handle))
        if rc != 0:
            msg = L.anuga_last_error(None)
            raise _native.DeviceError("anuga_ensemble_create failed (%d): %s"
                                      % (rc, msg.decode() if msg else ""))
        # per-scenario time state (the StepState of each)
        self._states = (_native.StepState * S)()
        for st in self._states:
            st.order = self.default_order
            st.recorded_min_timestep = self.evolve_max_timestep
            st.recorded_max_timestep = self.evolve_min_timestep
        self.number_of_steps = np.zeros(S, np.int64)
        self._meshes = [None] * S

    # -- handle --------------------------------------------------------------------
    def _call(self, name, *args):
        rc = getattr(_native.lib(), name)(self._handle, *args)
        _native.check(self._handle, rc, name)

    def close(self):
        """Release the device buffers now (also done on garbage collection)."""
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
        return self.size

    # -- set-up ----------------------------------------------------------------------
    def set_scenario(self, s, mesh, quantities, boundary, keep_mesh=False):
        """Upload scenario s: its mesh, centroid quantities {name: scalar, (N,)
        array or f(x, y) at the centroids} and boundaries {tag: kind}, where a
        kind is 'reflective', 'transmissive', 'transmissive_centroid',
        ('dirichlet', [stage, xmom, ymom]) or a boundary object of
        anuga_core_amd.boundaries."""
        self.set_mesh(s, mesh, keep=keep_mesh)
        for name, value in quantities.items():
            self.set_quantity(s, name, value, mesh=mesh)
        self.set_boundary(s, boundary, mesh=mesh)

    def set_mesh(self, s, mesh, keep=False):
        k = dict(nb=np.ascontiguousarray(mesh.neighbours, np.int64),
                 nbe=np.ascontiguousarray(mesh.neighbour_edges, np.int64),
                 sur=np.ascontiguousarray(mesh.surrogate_neighbours, np.int64),
                 nob=np.ascontiguousarray(mesh.number_of_boundaries, np.int64),
                 nrm=np.ascontiguousarray(mesh.normals), el=np.ascontiguousarray(mesh.edgelengths),
                 rad=np.ascontiguousarray(mesh.radii), area=np.ascontiguousarray(mesh.areas),
                 cen=np.ascontiguousarray(mesh.centroid_coordinates),
                 em=np.ascontiguousarray(mesh.edge_midpoint_coordinates))
        desc = _native.AnugaMesh(mesh.number_of_triangles, mesh.boundary_length,
                                 _ip(k["nb"]), _ip(k["nbe"]), _ip(k["sur"]), _ip(k["nob"]),
                                 _dptr(k["nrm"]), _dptr(k["el"]), _dptr(k["rad"]),
                                 _dptr(k["area"]), _dptr(k["cen"]), _dptr(k["em"]))
        self._call("anuga_ensemble_set_mesh", int(s), C.byref(desc))
        if keep:
            self._meshes[s] = mesh

    def set_quantity(self, s, name, value, mesh=None):
        """Centroid values of one quantity of scenario s."""
        if name not in ("stage", "xmomentum", "ymomentum", "elevation", "friction"):
            raise Exception("Unknown quantity %s" % name)
        N = int(self.number_of_triangles[s])
        if callable(value):
            mesh = mesh if mesh is not None else self._meshes[s]
            if mesh is None:
                raise Exception("a function value needs the scenario's mesh")
            cen = mesh.centroid_coordinates
            value = value(cen[:, 0], cen[:, 1])
        c = np.ascontiguousarray(np.broadcast_to(np.asarray(value, np.float64), (N,)))
        self._call("anuga_ensemble_set_centroid_values", int(s), _QIDS[name], _dptr(c))

    def set_boundary(self, s, boundary_map, mesh=None):
        mesh = mesh if mesh is not None else self._meshes[s]
        if mesh is None:
            raise Exception("set_boundary needs the scenario's mesh (its boundary tags)")
        tags = mesh.get_boundary_tags()
        for tag in tags:
            if tag not in boundary_map:
                raise Exception("Tag %s has not been bound to a boundary object" % tag)
        M = int(self.boundary_length[s])
        kind = np.zeros(max(M, 1), np.int32)
        vals = np.zeros((max(M, 1), 3))
        for tag, B in boundary_map.items():
            if tag not in tags:
                raise Exception("Tag %s provided does not exist in the scenario" % tag)
            ids = np.asarray(mesh.tag_boundary_cells[tag], np.int64)
            if isinstance(B, Boundary):
                code = B.code
                dv = getattr(B, "dirichlet_values", None)
            elif isinstance(B, tuple) and B[0] == "dirichlet":
                code, dv = KIND_DIRICHLET, np.asarray(B[1], np.float64)
            else:
                code, dv = _KINDS[B], None
            kind[ids] = code
            if code == KIND_DIRICHLET:
                vals[ids] = np.asarray(dv, np.float64)[:3]
        self._call("anuga_ensemble_set_boundary", int(s),
                   kind.ctypes.data_as(C.POINTER(C.c_int32)), _dptr(vals))

    # -- stepping --------------------------------------------------------------------------
    def _time_control(self, yieldtimes, finaltime):
        S = self.size
        tc = (_native.TimeControl * S)()
        inf = float("inf")
        for s in range(S):
            tc[s] = _native.TimeControl(self.CFL, self.evolve_max_timestep,
                                        self.evolve_min_timestep,
                                        inf if finaltime is None else float(finaltime),
                                        float(yieldtimes[s]), -1.0, EPSILON,
                                        self.max_smallsteps, self.default_order)
        return tc

    def _run(self, nsteps, tc):
        self._call("anuga_ensemble_evolve_steps", int(nsteps), tc, self._states)
        self.number_of_steps += np.array([st.steps_taken for st in self._states])
        return np.array([st.reached for st in self._states])

    def distribute_to_vertices_and_edges(self):
        """protect + extrapolation of every scenario, written back to the
        centroids (what Domain.distribute_to_vertices_and_edges does)."""
        me = np.empty(self.size)
        self._call("anuga_ensemble_distribute_to_vertices_and_edges", _dptr(me))
        self.mass_error = getattr(self, "mass_error", np.zeros(self.size)) + me

    def _reset_recorded(self):
        for st in self._states:
            st.recorded_min_timestep = self.evolve_max_timestep
            st.recorded_max_timestep = self.evolve_min_timestep

    def evolve_steps(self, nsteps, finaltime=None):
        """Exactly `nsteps` steps of every scenario (fewer for a scenario that
        reaches `finaltime`, relative to its start), with per-scenario
        update_timestep; one host synchronisation. The first call starts with
        distribute_to_vertices_and_edges, as Domain.evolve_steps does."""
        if not getattr(self, "evolved_called", False):
            self.distribute_to_vertices_and_edges()
            self.evolved_called = True
        for st in self._states:
            st.order = self.default_order
        self._reset_recorded()
        tc = self._time_control(np.full(self.size, np.inf), finaltime)
        return self._run(nsteps, tc)

    def evolve(self, yieldstep, finaltime, chunk=64):
        """Generator over common yield times (Domain.evolve for every
        scenario): each scenario steps to the next yield time and stops there,
        as _evolve_base does, while the others catch up; the ensemble then
        distributes and yields. Starts with a yield of the initial time."""
        self.distribute_to_vertices_and_edges()
        first = not getattr(self, "evolved_called", False)
        self.evolved_called = True
        for st in self._states:
            st.order = self.default_order
        yield_at = self.get_times().min() + float(yieldstep)
        if first:
            self.distribute_to_vertices_and_edges()
            yield self.get_times().min()
        while True:
            target = min(yield_at, float(finaltime))
            self._reset_recorded()
            tc = self._time_control(np.full(self.size, yield_at), finaltime)
            while True:
                reached = self._run(chunk, tc)
                if np.all(reached > 0):
                    break
            for s in np.nonzero(reached == 1)[0]:
                self._states[s].relative_time = float(finaltime)
            self.distribute_to_vertices_and_edges()
            yield target
            if target >= float(finaltime) - EPSILON:
                return
            yield_at += float(yieldstep)

    # -- results ------------------------------------------------------------------------
    def get_quantity(self, s, name):
        """Centroid values of a quantity of scenario s ('height' included)."""
        out = np.empty(int(self.number_of_triangles[s]))
        self._call("anuga_ensemble_get_centroid_values", int(s), _QIDS[name], _dptr(out))
        return out

    def get_time(self, s):
        return self._states[s].relative_time

    def get_times(self):
        return np.array([st.relative_time for st in self._states])

    def get_timestep(self, s):
        return self._states[s].timestep

    def get_water_volumes(self):
        v = np.empty(self.size)
        self._call("anuga_ensemble_get_water_volumes", _dptr(v))
        return v

    def get_boundary_flux_integrals(self):
        v = np.empty(self.size)
        self._call("anuga_ensemble_get_boundary_flux_integrals", _dptr(v))
        return v

    def synchronize(self):
        self._call("anuga_synchronize")

    def kernel_timing(self, name):
        ms = C.c_double()
        n = C.c_int64()
        self._call("anuga_kernel_timing", name.encode(), C.byref(ms), C.byref(n))
        return ms.value, n.value

    def reset_kernel_timing(self, enable=True):
        self._call("anuga_reset_kernel_timing", int(bool(enable)))

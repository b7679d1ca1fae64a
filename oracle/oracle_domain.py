"""Host orchestration of the oracle: a minimal restatement of the reference's
evolve loop for the DE flow algorithms, driving either CPU kernel set.

TEST INFRASTRUCTURE ONLY (imported by tests, the smoke check and the CPU
baseline of the benchmark; never by the product package).

Restates (paths relative to the anuga_core tree):
  * DE presets            anuga/shallow_water/shallow_water_domain.py:583-707,
                          config constants anuga/config.py
  * evolve / _evolve_base shallow_water_domain.py:2300-2407,
                          anuga/abstract_2d_finite_volumes/generic_domain.py:1715-1912
  * euler / rk2 steps     generic_domain.py:1914-2051
  * local flux updating   set_local_extrapolation_and_flux_updating
                          (shallow_water_domain.py:1193-1215) and the
                          compute_flux_update_frequency call of the euler step
                          (generic_domain.py:1960-1963)
  * update_timestep       generic_domain.py:2349-2415
  * update_boundary       generic_domain.py:2288-2306 with
                          Reflective_boundary.evaluate_segment (boundaries.py:235-292),
                          Transmissive_boundary / Dirichlet_boundary.evaluate_segment
                          (generic_boundary_conditions.py:173-264)
  * update_conserved_quantities shallow_water_domain.py:2089-2163
  * boundary flux integral anuga/operators/boundary_flux_integral_operator.py
  * get_water_volume      shallow_water_domain.py:1626-1653
  * Rate_operator         anuga/operators/rate_operators.py:149-260, applied
                          after every step (generic_domain.py:1849)
  * Inlet_operator        anuga/structures/inlet_operator.py:75-183 with
                          Inlet (anuga/structures/inlet.py:63-237)
  * Set_stage_operator    anuga/operators/set_stage_operator.py:25-58 applied
                          by Set_quantity.__call__ (set_quantity.py:95-131)
  * riverwall edge data   anuga/structures/riverwall.py:382-469 (the layout the
                          flux loop reads: edge_flux_type, per-rank elevation and
                          hydraulic row, edge_river_wall_counter)
Boundary objects are duck-typed: `.kind` in {'reflective', 'transmissive',
'dirichlet'} and, for Dirichlet, `.dirichlet_values`.
"""

from __future__ import annotations

import numpy as np

EPSILON = 1.0e-12          # anuga/config.py:12
G = 9.8                    # anuga/config.py:45
MAX_TIMESTEP = 1.0e+3      # anuga/config.py:154
MIN_TIMESTEP = 1.0e-6      # anuga/config.py:153
MAX_SMALLSTEPS = 50        # anuga/config.py:155


def _first(value):
    """float(value), or float(value[0]) for a vector (boundaries.py:502-506)."""
    try:
        return float(value)
    except TypeError:
        return float(value[0])


class OracleRateOperator:
    """Rate_operator.__call__ restated: indices (None: all triangles), rate
    (number, per-index array, or a function of t), factor (number or a
    function of t)."""

    def __init__(self, indices=None, rate=0.0, factor=1.0):
        self.indices = None if indices is None else np.asarray(indices, np.int64)
        self.rate = rate
        self.factor = factor
        self.cumulative_influx = 0.0

    def apply(self, d):
        t = d.get_time()
        dt = d.timestep
        rate = self.rate(t) if callable(self.rate) else self.rate
        rate = np.asarray(rate, np.float64)
        factor = float(self.factor(t)) if callable(self.factor) else self.factor
        idx = slice(None) if self.indices is None else self.indices
        areas = d.areas[idx]
        if np.all(rate >= 0.0):
            local_rates = factor * dt * rate
            self.cumulative_influx += float(np.sum(local_rates * areas))
            d.stage_c[idx] = d.stage_c[idx] + local_rates
        else:
            heights = d.stage_c[idx] - d.elev_c[idx]
            local_rates = np.maximum(factor * dt * rate, -heights)
            local_factors = np.where(local_rates < 0.0,
                                     (local_rates + heights) / (heights + 1.0e-10), 1.0)
            self.cumulative_influx += float(np.sum(local_rates * areas))
            d.stage_c[idx] = d.stage_c[idx] + local_rates
            d.xmom_c[idx] = d.xmom_c[idx] * local_factors
            d.ymom_c[idx] = d.ymom_c[idx] * local_factors


class OracleSetStageOperator:
    """Set_quantity.__call__ for 'stage' (anuga/operators/set_quantity.py:
    95-131, via Set_stage_operator, set_stage_operator.py:25-58) restated:
    indices (None: all triangles), stage a number or a function of t, of
    (x, y) or of (x, y, t) at the centroids, evaluated at the domain's time
    when the operator runs."""

    def __init__(self, indices=None, stage=0.0, centroids=None):
        self.indices = None if indices is None else np.asarray(indices, np.int64)
        self.stage = stage
        self.centroids = centroids
        self.cumulative_influx = 0.0

    def apply(self, d):
        t = d.get_time()
        idx = slice(None) if self.indices is None else self.indices
        v = self.stage
        if callable(v):
            import inspect
            npar = len(inspect.signature(v).parameters)
            if npar == 1:
                v = v(t)
            else:
                x, y = self.centroids[idx, 0], self.centroids[idx, 1]
                v = v(x, y) if npar == 2 else v(x, y, t)
        d.stage_c[idx] = v


class OracleInletOperator:
    """Inlet_operator.__call__ restated on a set of triangle indices."""

    def __init__(self, indices, Q=0.0, velocity=None, zero_velocity=False, default=0.0):
        self.idx = np.asarray(indices, np.int64)
        self.Q = Q
        self.velocity = velocity
        self.zero_velocity = zero_velocity
        self.default = default
        self.volume_integral = 0.0

    def _Q(self, t):
        return self.Q(t) if callable(self.Q) else self.Q

    def apply(self, d):
        idx = self.idx
        dt = d.timestep
        t = d.get_time()
        areas = d.areas.take(idx)
        area = np.sum(areas)
        stages = d.stage_c.take(idx)
        elev = d.elev_c.take(idx)
        depths = stages - elev
        current = np.sum(depths * areas)
        Q = 0.5 * (self._Q(t) + self._Q(t + dt))
        volume = Q * dt
        vp = 1.0e-6                        # anuga/config.py velocity_protection
        u = d.xmom_c.take(idx) * depths / (depths * depths + vp)
        v = d.ymom_c.take(idx) * depths / (depths * depths + vp)

        def momenta():
            dep = d.stage_c.take(idx) - d.elev_c.take(idx)
            if self.velocity is not None:
                d.xmom_c.put(idx, dep * self.velocity[0])
                d.ymom_c.put(idx, dep * self.velocity[1])
            else:
                d.xmom_c.put(idx, dep * u)
                d.ymom_c.put(idx, dep * v)
            if self.zero_velocity:
                d.xmom_c.put(idx, 0.0)
                d.ymom_c.put(idx, 0.0)

        if volume >= 0.0:
            # Inlet.set_stages_evenly (inlet.py:192-225)
            order = stages.argsort()
            summed_areas = np.cumsum(areas[order])
            summed_volume = np.zeros_like(areas)
            summed_volume[1:] = np.cumsum(summed_areas[:-1] * np.diff(stages[order]))
            index = np.nonzero(summed_volume <= volume)[0][-1]
            depth = (volume - summed_volume[index]) / summed_areas[index]
            stages[order[0:index + 1]] = stages[order[index]] + depth
            d.stage_c.put(idx, stages)
            self.volume_integral += volume
            momenta()
        elif current + volume >= 0.0:
            d.stage_c.put(idx, elev + (current + volume) / area)
            self.volume_integral += volume
            momenta()
        else:
            d.stage_c.put(idx, elev + 0.0)
            self.volume_integral -= current
            d.xmom_c.put(idx, 0.0)
            d.ymom_c.put(idx, 0.0)


class OracleDomain:
    def __init__(self, mesh, kernels, flow_algorithm="DE0"):
        self.mesh = mesh
        N = self.N = mesh.number_of_triangles
        M = self.M = mesh.boundary_length
        self.neighbours = np.ascontiguousarray(mesh.neighbours.ravel())
        self.neighbour_edges = np.ascontiguousarray(mesh.neighbour_edges.ravel())
        self.surrogate_neighbours = np.ascontiguousarray(mesh.surrogate_neighbours.ravel())
        self.number_of_boundaries = np.ascontiguousarray(mesh.number_of_boundaries)
        self.tri_full_flag = np.ones(N, np.int64)
        self.normals = np.ascontiguousarray(mesh.normals.ravel())
        self.edgelengths = np.ascontiguousarray(mesh.edgelengths.ravel())
        self.radii = np.ascontiguousarray(mesh.radii)
        self.areas = np.ascontiguousarray(mesh.areas)
        self.centroids = np.ascontiguousarray(mesh.centroid_coordinates.ravel())
        self.edge_midpoints = np.ascontiguousarray(mesh.edge_midpoint_coordinates.ravel())
        self.vertex_coordinates = np.ascontiguousarray(mesh.vertex_coordinates.ravel())
        z = lambda n: np.zeros(n)
        for q in ("stage", "xmom", "ymom", "elev", "height", "friction"):
            setattr(self, q + "_c", z(N))
            setattr(self, q + "_e", z(3 * N))
            setattr(self, q + "_v", z(3 * N))
        for q in ("stage", "xmom", "ymom"):
            setattr(self, q + "_b", z(max(M, 1)))
            setattr(self, q + "_ex", z(N))
            setattr(self, q + "_si", z(N))
            setattr(self, q + "_bk", z(N))
        self.x_centroid_work = z(N)
        self.y_centroid_work = z(N)
        self.max_speed = z(N)
        self.edge_timestep = z(3 * N) + 1.0e100
        self.boundary_flux_sum = z(3)
        self.boundary_flux_integral = 0.0
        self.operators = []
        self._no_riverwalls()

        self.g = G
        self.epsilon = EPSILON
        self.H0 = 1.0e-3
        self.low_froude = 0
        self.use_sloped_mannings = False
        self.extrapolate_velocity_second_order = 1
        self.evolve_max_timestep = MAX_TIMESTEP
        self.evolve_min_timestep = MIN_TIMESTEP
        self.max_smallsteps = MAX_SMALLSTEPS
        self.fixed_flux_timestep = None
        self.boundary_map = None
        self.max_flux_update_frequency = 1
        self.kernels = kernels
        kernels.bind(self)
        self.set_flow_algorithm(flow_algorithm)

        self.starttime = 0.0
        self.relative_time = 0.0
        self.evolved_called = False
        self.smallsteps = 0
        self.timestep = 0.0
        self.flux_timestep = 0.0
        self.dt_history = []
        self.mass_error = 0.0
        self.negative_cells = 0

    # -- riverwalls -----------------------------------------------------------
    def _no_riverwalls(self):
        N = self.N
        self.number_of_riverwall_edges = 0
        self.edge_flux_type = np.zeros(3 * N, np.int64)
        self.edge_river_wall_counter = np.zeros(3 * N, np.int64)
        self.riverwall_elevation = np.zeros(1)
        self.riverwall_rowIndex = np.zeros(1, np.int64)
        self.ncol_riverwall_hydraulic_properties = 5
        self.riverwall_hydraulic_properties = np.zeros(5)

    def set_riverwalls(self, edge_flux_type, riverwall_elevation, riverwall_rowIndex,
                       hydraulic_properties):
        """Riverwall data in the reference's layout: edge_flux_type (N*3, 1 on
        every half-edge of a wall), elevation and hydraulic-properties row per
        flagged half-edge in (triangle, edge) order, the hydraulic table
        (rows of Qfactor, s1, s2, h1, h2)."""
        eft = np.ascontiguousarray(edge_flux_type, np.int64).ravel()
        flagged = eft == 1
        R = int(flagged.sum())
        assert eft.shape == (3 * self.N,)
        assert len(riverwall_elevation) == R and len(riverwall_rowIndex) == R
        hp = np.ascontiguousarray(hydraulic_properties, np.float64)
        self.number_of_riverwall_edges = R
        self.edge_flux_type = eft
        self.edge_river_wall_counter = np.where(flagged, np.cumsum(flagged), 0).astype(np.int64)
        self.riverwall_elevation = np.ascontiguousarray(riverwall_elevation, np.float64)
        self.riverwall_rowIndex = np.ascontiguousarray(riverwall_rowIndex, np.int64)
        self.ncol_riverwall_hydraulic_properties = hp.shape[1]
        self.riverwall_hydraulic_properties = hp.ravel().copy()
        self.kernels.bind(self)

    # -- presets ------------------------------------------------------------
    # shallow_water_domain.py:583-895: CFL, timestepping, minimum_allowed_height,
    # beta, beta_dry
    _PRESETS = {"DE0": (0.9, "euler", 1.0e-12, 0.5, 0.0),
                "DE1": (1.0, "rk2", 1.0e-5, 1.0, 0.0),
                "DE2": (1.0, "rk3", 1.0e-5, 1.0, 0.0),
                "DE1_7": (1.0, "rk2", 1.0e-12, 0.75, 0.1),
                "DE0_7": (0.9, "euler", 1.0e-12, 0.7, 0.1)}

    def set_flow_algorithm(self, algorithm="DE0"):
        algorithm = str(algorithm).replace(".", "_")
        if algorithm not in self._PRESETS:
            raise ValueError("flow algorithm %s is not supported by the oracle" % algorithm)
        cfl, method, mah, b, bd = self._PRESETS[algorithm]
        self.CFL = cfl
        self.timestepping_method = method
        self.minimum_allowed_height = mah
        self.flow_algorithm = algorithm
        self.beta_w = self.beta_uh = self.beta_vh = b
        self.beta_w_dry = self.beta_uh_dry = self.beta_vh_dry = bd
        self.default_order = 2

    # -- quantities -----------------------------------------------------------
    _NAMES = {"stage": "stage", "xmomentum": "xmom", "ymomentum": "ymom",
              "elevation": "elev", "friction": "friction", "height": "height"}

    def set_centroid_values(self, name, values):
        getattr(self, self._NAMES[name] + "_c")[:] = values

    def set_vertex_values(self, name, vertex_values):
        """(N,3) vertex values -> vertex, edge and centroid values
        (quantity.c:665-688, interpolate)."""
        q = self._NAMES[name]
        v = np.asarray(vertex_values, np.float64).reshape(self.N, 3)
        getattr(self, q + "_v")[:] = v.ravel()
        q0, q1, q2 = v[:, 0], v[:, 1], v[:, 2]
        getattr(self, q + "_c")[:] = (q0 + q1 + q2) / 3.0
        e = getattr(self, q + "_e").reshape(self.N, 3)
        e[:, 0] = 0.5 * (q1 + q2)
        e[:, 1] = 0.5 * (q0 + q2)
        e[:, 2] = 0.5 * (q0 + q1)

    def get_centroid_values(self, name):
        return getattr(self, self._NAMES[name] + "_c").copy()

    def set_boundary(self, boundary_map):
        self.boundary_map = dict(boundary_map)

    # -- boundary conditions ---------------------------------------------------
    def update_boundary(self):
        m = self.mesh
        nrm = self.normals.reshape(self.N, 6)
        for tag, ids in m.tag_boundary_cells.items():
            B = self.boundary_map.get(tag) if self.boundary_map else None
            if B is None:
                continue
            ids = np.asarray(ids, np.int64)
            vol = m.boundary_cells[ids]
            edge = m.boundary_edges[ids]
            se = self.stage_e.reshape(self.N, 3)
            xe = self.xmom_e.reshape(self.N, 3)
            ye = self.ymom_e.reshape(self.N, 3)
            if B.kind == "reflective":
                n1 = nrm[vol, 2 * edge]
                n2 = nrm[vol, 2 * edge + 1]
                self.stage_b[ids] = se[vol, edge]
                q1 = xe[vol, edge]
                q2 = ye[vol, edge]
                r1 = -q1 * n1 - q2 * n2
                r2 = -q1 * n2 + q2 * n1
                self.xmom_b[ids] = n1 * r1 - n2 * r2
                self.ymom_b[ids] = n2 * r1 + n1 * r2
            elif B.kind == "transmissive":
                self.stage_b[ids] = se[vol, edge]
                self.xmom_b[ids] = xe[vol, edge]
                self.ymom_b[ids] = ye[vol, edge]
            elif B.kind == "dirichlet":
                v = np.asarray(B.dirichlet_values, np.float64)
                self.stage_b[ids] = v[0]
                self.xmom_b[ids] = v[1]
                self.ymom_b[ids] = v[2]
            elif B.kind == "time":
                # Time_boundary.evaluate_segment (generic_boundary_conditions.py:383-417)
                v = np.asarray(B.function(self.get_time()), np.float64)
                self.stage_b[ids] = v[0]
                self.xmom_b[ids] = v[1]
                self.ymom_b[ids] = v[2]
            elif B.kind == "time_stage_zero_momentum":
                # boundaries.py:620-636
                self.stage_b[ids] = _first(B.function(self.get_time()))
                self.xmom_b[ids] = 0.0
                self.ymom_b[ids] = 0.0
            elif B.kind == "transmissive_n_momentum_zero_t_momentum_set_stage":
                # boundaries.py:474-517
                n1 = nrm[vol, 2 * edge]
                n2 = nrm[vol, 2 * edge + 1]
                self.stage_b[ids] = _first(B.function(self.get_time()))
                ndotq = n1 * xe[vol, edge] + n2 * ye[vol, edge]
                self.xmom_b[ids] = ndotq * n1
                self.ymom_b[ids] = ndotq * n2
            elif B.kind == "transmissive_momentum_set_stage":
                # boundaries.py:344-373 (time rolled over past the function's range)
                t = self.get_time()
                if hasattr(B.function, "time"):
                    while t > B.function.time[-1]:
                        t -= B.function.time[-1]
                self.stage_b[ids] = _first(B.function(t))
                self.xmom_b[ids] = xe[vol, edge]
                self.ymom_b[ids] = ye[vol, edge]
            elif B.kind == "transmissive_stage_zero_momentum":
                # boundaries.py:542-550
                self.stage_b[ids] = se[vol, edge]
                self.xmom_b[ids] = 0.0
                self.ymom_b[ids] = 0.0
            elif B.kind in ("file", "field"):
                # File_boundary / Field_boundary: per-edge values at the
                # domain's time (generic_boundary_conditions.py:646-680)
                v = B.values_at(self.get_time(), ids)
                self.stage_b[ids] = v[:, 0]
                self.xmom_b[ids] = v[:, 1]
                self.ymom_b[ids] = v[:, 2]
            else:
                raise ValueError("unsupported boundary kind %r" % B.kind)

    # -- step pieces --------------------------------------------------------------
    def distribute_to_vertices_and_edges(self):
        self.mass_error += self.kernels.protect()
        if self.kernels.extrapolate() == -1:
            raise RuntimeError("Internal neighbour not found")

    def compute_fluxes(self, substep):
        self.flux_timestep = self.kernels.compute_fluxes(substep, self.evolve_max_timestep)

    def compute_forcing_terms(self):
        self.kernels.friction()

    def update_conserved_quantities(self):
        self.kernels.update(self.timestep)
        self.negative_cells += self.kernels.fix_negative()

    def update_timestep(self, yieldstep, finaltime):
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
                    raise Exception("WARNING: Too small timestep %.16f reached even after "
                                    "%d steps of 1 order scheme" % (timestep, self.max_smallsteps))
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

    def evolve_one_euler_step(self, yieldstep, finaltime):
        """generic_domain.py:1914-1972"""
        self.distribute_to_vertices_and_edges()
        self.update_boundary()
        self.compute_fluxes(0)
        self.compute_forcing_terms()
        self.update_timestep(yieldstep, finaltime)
        if self.max_flux_update_frequency != 1:
            self.kernels.compute_flux_update_frequency(self.timestep)
        self.update_conserved_quantities()

    def set_local_extrapolation_and_flux_updating(self, nlevels=8):
        """shallow_water_domain.py:1193-1215"""
        self.max_flux_update_frequency = 2 ** nlevels
        if self.max_flux_update_frequency != 1:
            if self.timestepping_method != "euler":
                raise Exception("Local extrapolation and flux updating only supported with "
                                "euler timestepping")
        self.kernels.set_max_flux_update_frequency(self.max_flux_update_frequency)

    def evolve_one_rk2_step(self, yieldstep, finaltime):
        self.kernels.backup()
        self.distribute_to_vertices_and_edges()
        self.update_boundary()
        self.compute_fluxes(0)
        self.compute_forcing_terms()
        self.update_timestep(yieldstep, finaltime)
        self.update_conserved_quantities()
        self.relative_time = self.relative_time + self.timestep
        self.distribute_to_vertices_and_edges()
        self.update_boundary()
        self.compute_fluxes(1)
        self.compute_forcing_terms()
        self.update_conserved_quantities()
        self.kernels.saxpy(0.5, 0.5)

    def evolve_one_rk3_step(self, yieldstep, finaltime):
        """generic_domain.py:2053-2179"""
        self.kernels.backup()
        initial = self.relative_time
        self.distribute_to_vertices_and_edges()
        self.update_boundary()
        self.compute_fluxes(0)
        self.compute_forcing_terms()
        self.update_timestep(yieldstep, finaltime)
        self.update_conserved_quantities()
        self.relative_time = self.relative_time + self.timestep
        self.distribute_to_vertices_and_edges()
        self.update_boundary()
        self.compute_fluxes(1)
        self.compute_forcing_terms()
        self.update_conserved_quantities()
        self.kernels.saxpy(0.25, 0.75)
        self.relative_time = initial + self.timestep * 0.5
        self.distribute_to_vertices_and_edges()
        self.update_boundary()
        self.compute_fluxes(2)
        self.compute_forcing_terms()
        self.update_conserved_quantities()
        self.kernels.saxpy(2.0, 1.0)
        for q in (self.stage_c, self.xmom_c, self.ymom_c):
            q[:] = q / 3.0
        self.relative_time = initial + self.timestep

    def _apply_boundary_flux_integral(self):
        """boundary_flux_integral_operator.py:46-62"""
        dt = self.timestep
        s = self.boundary_flux_sum
        if self.timestepping_method == "euler":
            self.boundary_flux_integral = self.boundary_flux_integral + dt * s[0]
        elif self.timestepping_method == "rk2":
            self.boundary_flux_integral = self.boundary_flux_integral + 0.5 * dt * s[0:2].sum()
        else:
            self.boundary_flux_integral = (self.boundary_flux_integral
                                           + 1.0 / 6.0 * dt * (s[0] + s[1] + 4.0 * s[2]))
        s[:] = 0.0

    # -- evolve ---------------------------------------------------------------------
    def get_time(self):
        return self.starttime + self.relative_time

    def set_starttime(self, time):
        """generic_domain.py:1607-1615"""
        if self.evolved_called:
            raise Exception("Can't change simulation start time once evolve has been called")
        self.starttime = float(time)
        # starttime is now the origin of relative_time
        self.relative_time = 0.0

    def evolve(self, yieldstep=None, finaltime=None, max_steps=None):
        """Generator over yield times; `max_steps` stops after that many steps
        (a test convenience, not in the reference)."""
        if self.boundary_map is None:
            raise AssertionError("Boundary tags must be bound to boundary objects before evolving")
        self.distribute_to_vertices_and_edges()
        skip_initial_step = self.evolved_called
        self.evolved_called = True
        yieldstep = self.evolve_max_timestep if yieldstep is None else float(yieldstep)
        self._order_ = self.default_order
        self.relative_finaltime = float(finaltime) - self.starttime
        if self.relative_finaltime < self.relative_time:
            return
        self.relative_yieldtime = self.relative_time + yieldstep
        self.recorded_min_timestep = self.evolve_max_timestep
        self.recorded_max_timestep = self.evolve_min_timestep
        self.number_of_steps = 0
        if not skip_initial_step:
            self.distribute_to_vertices_and_edges()
            self.update_boundary()
            yield self.get_time()
        steps = 0
        while True:
            initial = self.relative_time
            if self.timestepping_method == "euler":
                self.evolve_one_euler_step(yieldstep, self.relative_finaltime)
            elif self.timestepping_method == "rk2":
                self.evolve_one_rk2_step(yieldstep, self.relative_finaltime)
            else:
                self.evolve_one_rk3_step(yieldstep, self.relative_finaltime)
            self._apply_boundary_flux_integral()
            for op in self.operators:
                op.apply(self)
            self.relative_time = initial + self.timestep
            self.dt_history.append(self.timestep)
            self.number_of_steps += 1
            steps += 1
            if self.relative_time >= self.relative_finaltime - EPSILON:
                if self.relative_time > self.relative_finaltime:
                    raise Exception("WARNING (domain.py): time overshot finaltime. ")
                self.relative_time = self.relative_finaltime
                self.distribute_to_vertices_and_edges()
                self.update_boundary()
                yield self.get_time()
                break
            if max_steps is not None and steps >= max_steps:
                break
            if self.relative_time >= self.relative_yieldtime:
                self.distribute_to_vertices_and_edges()
                self.update_boundary()
                yield self.get_time()
                self.relative_yieldtime += yieldstep
                self.recorded_min_timestep = self.evolve_max_timestep
                self.recorded_max_timestep = self.evolve_min_timestep
                self.number_of_steps = 0

    def get_water_volume(self):
        return float(np.sum(self.areas * self.height_c))

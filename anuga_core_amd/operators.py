"""Fractional-step operators applied on the device after every step.

Rate_operator (anuga/operators/rate_operators.py) adds water over a set of
triangles at a rate in m/s (rain, or extraction with negative rates): after
each step the stage rises by factor * dt * rate; where a rate is negative the
loss is capped at the depth and the momenta scale with the depth that
remains. The operator's triangles come from `indices`, a `polygon`, a circle
(`center`, `radius`) or a Region-like object with `indices` (all triangles
when none is given).

Static rates (a number, a per-triangle array or a function of (x, y)) live on
the device and are applied inside the captured step. A rate or factor that is
a function of time is evaluated on the host after each step at the time the
reference uses (the step's start for euler, start + dt for rk2 and rk3), and
the domain then takes host-sequenced steps.
"""

from __future__ import annotations

import ctypes as C
import inspect

import numpy as np


def _npar(f):
    try:
        return len(inspect.signature(f).parameters)
    except (TypeError, ValueError):
        return 1


def _inside_polygon(points, polygon):
    """Even-odd rule; points on an edge count as inside."""
    poly = np.asarray(polygon, np.float64)
    x, y = points[:, 0], points[:, 1]
    inside = np.zeros(len(points), bool)
    n = len(poly)
    for i in range(n):
        x0, y0 = poly[i]
        x1, y1 = poly[(i + 1) % n]
        cross = ((y0 > y) != (y1 > y))
        with np.errstate(divide="ignore", invalid="ignore"):
            xi = x0 + (y - y0) * (x1 - x0) / (y1 - y0)
        inside ^= cross & (x < xi)
        # on the edge
        dx, dy = x1 - x0, y1 - y0
        t = ((x - x0) * dx + (y - y0) * dy) / max(dx * dx + dy * dy, 1e-300)
        on = (np.abs((x - x0) * dy - (y - y0) * dx) <= 1e-12 * max(abs(dx) + abs(dy), 1.0)) & \
             (t >= 0) & (t <= 1)
        inside |= on
    return inside


def _line_triangles(mesh, line, samples=400):
    """Triangles crossed by a segment (those holding points sampled along it)."""
    (x0, y0), (x1, y1) = np.asarray(line, np.float64)
    t = np.linspace(0.0, 1.0, samples + 1)
    pts = np.stack([x0 + t * (x1 - x0), y0 + t * (y1 - y0)], axis=1)
    tri = mesh.find_triangles(pts)
    return np.unique(tri[tri >= 0])


class Rate_operator:
    """Rate_operator(domain, rate=0.0, factor=1.0, region=None, indices=None,
    polygon=None, center=None, radius=None, default_rate=0.0, ...)."""

    def __init__(self, domain, rate=0.0, factor=1.0, region=None, indices=None, polygon=None,
                 center=None, radius=None, default_rate=0.0, description=None, label=None,
                 logging=False, verbose=False, monitor=False):
        self.domain = domain
        mesh = domain.mesh
        cen = mesh.centroid_coordinates
        if region is not None:
            indices = getattr(region, "indices", region)
        elif polygon is not None:
            indices = np.nonzero(_inside_polygon(cen, polygon))[0]
        elif center is not None and radius is not None:
            c = np.asarray(center, np.float64)
            indices = np.nonzero(((cen - c) ** 2).sum(axis=1) <= radius * radius)[0]
        self.indices = None if indices is None else np.unique(np.asarray(indices, np.int64))
        self.rate = rate
        self.factor = factor
        self.default_rate = default_rate
        self.cumulative_influx = 0.0
        sel = slice(None) if self.indices is None else self.indices
        self._coord = cen[sel]
        # static or time-dependent
        self.time_dependent = (callable(rate) and _npar(rate) in (1, 3)) or callable(factor)
        rates, scalar = self._rate_values(0.0)
        f = self._factor(0.0)
        n = 0 if self.indices is None else len(self.indices)
        idx = None if self.indices is None else np.ascontiguousarray(self.indices)
        self._id = C.c_int64()
        domain._call("anuga_add_rate_operator", n,
                     None if idx is None else idx.ctypes.data_as(C.POINTER(C.c_int64)),
                     None if rates is None else rates.ctypes.data_as(C.POINTER(C.c_double)),
                     C.c_double(scalar), C.c_double(f), C.byref(self._id))
        self._keep = (idx, rates)
        domain.fractional_step_operators.append(self)

    # -- values ---------------------------------------------------------------
    def _rate_values(self, t):
        """(per-triangle rates or None, scalar rate) at time t."""
        r = self.rate
        if callable(r):
            npar = _npar(r)
            x, y = self._coord[:, 0], self._coord[:, 1]
            if npar == 2:
                return np.ascontiguousarray(np.broadcast_to(r(x, y), (len(x),)), np.float64), 0.0
            if npar == 3:
                return np.ascontiguousarray(np.broadcast_to(r(x, y, t), (len(x),)),
                                            np.float64), 0.0
            try:
                v = r(t)
            except Exception:
                v = self.default_rate
            return None, float(v)
        a = np.asarray(r, np.float64)
        if a.ndim == 0:
            return None, float(a)
        if self.indices is not None and len(a) == self.domain.number_of_elements:
            a = a[self.indices]
        return np.ascontiguousarray(a, np.float64), 0.0

    def _factor(self, t):
        f = self.factor
        return float(f(t)) if callable(f) else float(f)

    def update(self, t, timestep=None):
        """Re-evaluate a time-dependent rate or factor at time t."""
        rates, scalar = self._rate_values(t)
        self._keep = (self._keep[0], rates)
        self.domain._call("anuga_update_rate_operator", self._id.value,
                          None if rates is None else rates.ctypes.data_as(C.POINTER(C.c_double)),
                          C.c_double(scalar), C.c_double(self._factor(t)))

    def get_cumulative_influx(self):
        v = C.c_double()
        self.domain._call("anuga_get_operator_influx", self._id.value, C.byref(v))
        return v.value

    def __repr__(self):
        return "Rate_operator(%s)" % ("all triangles" if self.indices is None
                                      else "%d triangles" % len(self.indices))


def _region_indices(domain, region=None, indices=None, polygon=None, center=None, radius=None,
                    line=None):
    """Triangle indices of a region as anuga.Region selects them (None: all)."""
    mesh = domain.mesh
    cen = mesh.centroid_coordinates
    if region is not None:
        indices = getattr(region, "indices", region)
    elif polygon is not None:
        indices = np.nonzero(_inside_polygon(cen, polygon))[0]
    elif center is not None and radius is not None:
        c = np.asarray(center, np.float64)
        indices = np.nonzero(((cen - c) ** 2).sum(axis=1) <= radius * radius)[0]
    elif line is not None:
        indices = _line_triangles(mesh, line)
    return None if indices is None else np.unique(np.asarray(indices, np.int64))


class Set_stage_operator:
    """Set_stage_operator(domain, stage=None, region=None, indices=None,
    polygon=None, center=None, radius=None, line=None, ...)
    (anuga/operators/set_stage_operator.py:25-58): after every step the stage
    centroid values of the region become `stage`, a number or a function of
    t, of (x, y) or of (x, y, t) at the centroids (Set_quantity.get_value,
    anuga/operators/set_quantity.py:133-158); momenta are unchanged.
    `indices` None: every triangle; []: none. Static values are applied on
    the device inside the captured step; a value that depends on t is
    evaluated on the host at the domain's time after each step, and the
    domain then takes host-sequenced steps. stage None: the operator does
    nothing."""

    def __init__(self, domain, stage=None, region=None, indices=None, polygon=None, center=None,
                 radius=None, line=None, description=None, label=None, logging=False,
                 verbose=False):
        self.domain = domain
        self.indices = _region_indices(domain, region, indices, polygon, center, radius, line)
        self.center, self.radius = center, radius
        cen = domain.mesh.centroid_coordinates
        sel = slice(None) if self.indices is None else self.indices
        self._coord = cen[sel]
        self.stage = stage
        self.time_dependent = callable(stage) and _npar(stage) in (1, 3)
        self._id = None
        if stage is None:
            return
        values, scalar = self._values(domain.get_time())
        n = 0 if self.indices is None else len(self.indices)
        idx = None if self.indices is None else np.ascontiguousarray(self.indices)
        self._id = C.c_int64()
        domain._call("anuga_add_set_stage_operator", n,
                     None if idx is None else idx.ctypes.data_as(C.POINTER(C.c_int64)),
                     None if values is None else values.ctypes.data_as(C.POINTER(C.c_double)),
                     C.c_double(scalar), C.byref(self._id))
        self._keep = (idx, values)
        domain.fractional_step_operators.append(self)

    def _values(self, t):
        """(per-triangle stages or None, scalar stage) at time t."""
        v = self.stage
        if v is None:
            return None, 0.0
        if callable(v):
            x, y = self._coord[:, 0], self._coord[:, 1]
            npar = _npar(v)
            if npar == 1:
                return None, float(v(t))
            r = v(x, y) if npar == 2 else v(x, y, t)
            return np.ascontiguousarray(np.broadcast_to(r, (len(x),)), np.float64), 0.0
        return None, float(v)

    def get_stage(self, t=None):
        return self._values(self.domain.get_time() if t is None else t)

    def update(self, t, timestep=None):
        """Re-evaluate a time-dependent stage at time t."""
        values, scalar = self._values(t)
        self._keep = (self._keep[0], values)
        self.domain._call("anuga_update_rate_operator", self._id.value,
                          None if values is None else values.ctypes.data_as(C.POINTER(C.c_double)),
                          C.c_double(scalar), C.c_double(1.0))

    def get_cumulative_influx(self):
        return 0.0

    def __repr__(self):
        return "Set_stage_operator(%s)" % ("all triangles" if self.indices is None
                                           else "%d triangles" % len(self.indices))


class Inlet_operator:
    """Inlet_operator(domain, region, Q=0.0, velocity=None, zero_velocity=False,
    default=0.0, ...) (anuga/structures/inlet_operator.py): a discharge Q
    (m^3/s, a number or a function of t) into a region after every step. A
    positive volume fills the region's lowest stages first to a level
    surface (Inlet.set_stages_evenly, anuga/structures/inlet.py:192-225); a
    negative one lowers the region to a uniform depth, or empties it. The
    momenta follow the region's velocities (or `velocity`, or zero).

    Regions of up to 2048 triangles are levelled on the device, one
    workgroup per inlet after every step (anuga_add_inlet_operator), with
    numpy's summation order and the stages sorted with equal stages in
    region order: the same bits as the reference's numpy wherever equal
    stages have equal areas (otherwise the cumulative area may round in
    another order). A constant Q keeps runs of steps on the device; a Q(t)
    is evaluated on the host each step (two scalars uploaded). Larger
    regions (or device=False) apply the reference's arithmetic on the host
    to the region's state, read and written back each step. `region`:
    triangle indices, an object with `indices`, a line, or a polygon
    (triangles whose centroid lies inside)."""

    velocity_protection = 1.0e-6    # anuga/config.py velocity_protection
    DEVICE_MAX = 2048

    def __init__(self, domain, region, Q=0.0, velocity=None, zero_velocity=False, default=0.0,
                 description=None, label=None, logging=False, verbose=False, device=True):
        self.domain = domain
        if hasattr(region, "indices"):
            idx = region.indices
        else:
            a = np.asarray(region)
            if a.ndim == 2 and a.shape == (2, 2) and a.dtype.kind == "f":
                idx = _line_triangles(domain.mesh, a)
            elif a.ndim == 2 and a.shape[1] == 2 and a.dtype.kind == "f":
                idx = np.nonzero(_inside_polygon(domain.mesh.centroid_coordinates, a))[0]
            else:
                idx = a
        self.triangle_indices = np.ascontiguousarray(idx, np.int64)
        if len(self.triangle_indices) == 0:
            raise Exception("No triangles have been identified in region")
        self.areas = domain.mesh.areas.take(self.triangle_indices)
        self.area = np.sum(self.areas)
        self.Q = Q
        if velocity is not None:
            assert len(velocity) == 2
        self.velocity = velocity
        self.zero_velocity = zero_velocity
        self.default = default
        self.applied_Q = 0.0
        self.total_applied_volume = 0.0
        self.total_requested_volume = 0.0
        self.volume_integral = 0.0
        self.host_side = not (device and len(self.triangle_indices) <= self.DEVICE_MAX)
        self.time_dependent = self.host_side or callable(Q)
        if not self.host_side:
            vmode = 2 if zero_velocity else (1 if velocity is not None else 0)
            u, v = (float(velocity[0]), float(velocity[1])) if velocity is not None else (0.0, 0.0)
            q0 = 0.0 if callable(Q) else float(Q)
            self._id = C.c_int64()
            areas = np.ascontiguousarray(self.areas, np.float64)
            domain._call("anuga_add_inlet_operator", len(self.triangle_indices),
                         self.triangle_indices.ctypes.data_as(C.POINTER(C.c_int64)),
                         areas.ctypes.data_as(C.POINTER(C.c_double)), C.c_double(self.area),
                         vmode, C.c_double(u), C.c_double(v), C.c_double(q0), C.c_double(q0),
                         C.byref(self._id))
        domain.fractional_step_operators.append(self)

    def update(self, t, timestep=0.0):
        """Q at the operator's time and a timestep later, for the device."""
        domain = self.domain
        Q1 = self.update_Q(t)
        Q2 = self.update_Q(t + timestep)
        domain._call("anuga_update_inlet_operator", self._id.value, C.c_double(Q1),
                     C.c_double(Q2))

    def _stats(self):
        st = np.zeros(4)
        self.domain._call("anuga_get_inlet_operator_stats", self._id.value,
                          st.ctypes.data_as(C.POINTER(C.c_double)))
        return st

    def update_Q(self, t):
        if callable(self.Q):
            try:
                return self.Q(t)
            except Exception as e:
                if type(e).__name__ in ("Modeltime_too_early", "Modeltime_too_late"):
                    return self.default(t) if callable(self.default) else self.default
                raise
        return self.Q

    def apply(self, t, timestep):
        d = self.domain
        idx = self.triangle_indices
        n = len(idx)
        st = np.empty((n, 4))
        ip = idx.ctypes.data_as(C.POINTER(C.c_int64))
        d._call("anuga_get_centroid_subset", n, ip, st.ctypes.data_as(C.POINTER(C.c_double)))
        stages, xmoms, ymoms, elev = st[:, 0].copy(), st[:, 1].copy(), st[:, 2].copy(), st[:, 3]
        depths = stages - elev
        current_volume = np.sum(depths * self.areas)
        assert current_volume >= 0.0
        Q1 = self.update_Q(t)
        Q2 = self.update_Q(t + timestep)
        Q = 0.5 * (Q1 + Q2)
        volume = Q * timestep
        self.applied_Q = Q
        vp = self.velocity_protection
        u = xmoms * depths / (depths * depths + vp)
        v = ymoms * depths / (depths * depths + vp)
        if volume >= 0.0:
            # set_stages_evenly
            order = stages.argsort()
            summed_areas = np.cumsum(self.areas[order])
            summed_volume = np.zeros_like(self.areas)
            summed_volume[1:] = np.cumsum(summed_areas[:-1] * np.diff(stages[order]))
            index = np.nonzero(summed_volume <= volume)[0][-1]
            depth = (volume - summed_volume[index]) / summed_areas[index]
            stages[order[0:index + 1]] = stages[order[index]] + depth
            self.volume_integral += volume
            self.total_requested_volume += volume
            depths = stages - elev
            xmoms, ymoms = self._momenta(depths, u, v)
        elif current_volume + volume >= 0.0:
            depth = (current_volume + volume) / self.area
            stages = elev + depth
            self.total_requested_volume += volume
            self.volume_integral += volume
            depths = stages - elev
            xmoms, ymoms = self._momenta(depths, u, v)
        else:
            stages = elev + 0.0
            self.total_requested_volume += volume
            volume = -current_volume
            self.applied_Q = -current_volume / timestep
            self.volume_integral -= current_volume
            xmoms = np.zeros(n)
            ymoms = np.zeros(n)
        self.total_applied_volume += volume
        out = np.ascontiguousarray(np.stack([stages, xmoms, ymoms, elev], axis=1))
        d._call("anuga_set_centroid_subset", n, ip, out.ctypes.data_as(C.POINTER(C.c_double)))

    def _momenta(self, depths, u, v):
        if self.zero_velocity:
            return np.zeros_like(depths), np.zeros_like(depths)
        if self.velocity is not None:
            return depths * self.velocity[0], depths * self.velocity[1]
        return depths * u, depths * v

    def get_cumulative_influx(self):
        if not self.host_side:
            st = self._stats()
            (self.volume_integral, self.total_applied_volume, self.total_requested_volume,
             self.applied_Q) = (float(x) for x in st)
        return self.volume_integral

    def get_total_applied_volume(self):
        self.get_cumulative_influx()
        return self.total_applied_volume

    def get_total_requested_volume(self):
        self.get_cumulative_influx()
        return self.total_requested_volume

    def get_applied_Q(self):
        self.get_cumulative_influx()
        return self.applied_Q

    def __repr__(self):
        return "Inlet_operator(%d triangles)" % len(self.triangle_indices)

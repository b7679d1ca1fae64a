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

    def update(self, t):
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

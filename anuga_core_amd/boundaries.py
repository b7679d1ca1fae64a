"""Boundary condition objects of the DE step.

They carry no per-step host work: the flux kernel evaluates each boundary edge's
right-hand state on the device from the boundary triangle's own edge values,
which is exactly what the reference computes on the host between extrapolation
and the flux call:

* Reflective_boundary  anuga/shallow_water/boundaries.py:235-292 (stage copied,
  momentum rotated into the edge frame, normal component negated, rotated back)
* Transmissive_boundary anuga/abstract_2d_finite_volumes/generic_boundary_conditions.py:139-195
  (edge values, or centroid values with domain.centroid_transmissive_bc)
* Dirichlet_boundary   generic_boundary_conditions.py:198-264 (constant stage,
  xmomentum, ymomentum)
* Transmissive_stage_zero_momentum_boundary  boundaries.py:520-550

Time-dependent boundaries (Time_boundary, Time_stage_zero_momentum_boundary,
Transmissive_n_momentum_zero_t_momentum_set_stage_boundary,
Transmissive_momentum_set_stage_boundary) evaluate their Python function on
the host at each substep time and upload the values; the kernel applies them.
"""

from __future__ import annotations

import numpy as np

# boundary kind codes shared with the device library (include/anuga_hip.h)
KIND_REFLECTIVE = 0
KIND_TRANSMISSIVE = 1
KIND_DIRICHLET = 2
KIND_TRANSMISSIVE_CENTROID = 3


class Boundary:
    kind = None

    def __repr__(self):
        return "%s()" % type(self).__name__


class Reflective_boundary(Boundary):
    """Reflective boundary: stage copied, momentum reflected in the edge normal."""
    kind = "reflective"
    code = KIND_REFLECTIVE

    def __init__(self, domain=None):
        if domain is None:
            raise Exception("Domain must be specified for reflective boundary")
        self.domain = domain


class Transmissive_boundary(Boundary):
    """Transmissive boundary: the boundary state equals the interior edge state."""
    kind = "transmissive"
    code = KIND_TRANSMISSIVE

    def __init__(self, domain=None):
        if domain is None:
            raise Exception("Domain must be specified for transmissive boundary")
        self.domain = domain


class Dirichlet_boundary(Boundary):
    """Dirichlet boundary: constant [stage, xmomentum, ymomentum]."""
    kind = "dirichlet"
    code = KIND_DIRICHLET

    def __init__(self, dirichlet_values=None):
        if dirichlet_values is None:
            raise Exception("Must specify one value for each quantity")
        v = np.array(dirichlet_values, dtype=np.float64).ravel()
        if v.size < 3:
            raise Exception("Dirichlet_boundary needs stage, xmomentum and ymomentum values")
        self.dirichlet_values = v

    def __repr__(self):
        return "Dirichlet boundary (%s)" % self.dirichlet_values


KIND_SET_STAGE_NORMAL_MOMENTUM = 4
KIND_TRANSMISSIVE_STAGE_ZERO_MOMENTUM = 5
KIND_SET_STAGE_TRANSMISSIVE_MOMENTUM = 6


class _TimeFunctionBoundary(Boundary):
    """A boundary whose values come from function(t) at the domain's time
    (Boundary.get_boundary_values, generic_boundary_conditions.py:95-136:
    default_boundary replaces the values once the function reports a model
    time beyond its range). The domain evaluates it at every substep time on
    the host and uploads the values; the flux kernel applies them."""
    time_dependent = True

    def __init__(self, domain=None, function=None, default_boundary=None, verbose=False):
        if domain is None:
            raise Exception("Domain must be specified for this type of boundary")
        if function is None:
            raise Exception("Function must be specified for this type of boundary")
        self.domain = domain
        self.function = function
        self.default_boundary = default_boundary
        self.default_boundary_invoked = False
        self.verbose = verbose

    def get_time(self):
        return self.domain.get_time()

    def get_boundary_values(self, t=None):
        if t is None:
            t = self.get_time()
        try:
            return self.function(t)
        except Exception as e:
            if self.default_boundary is None or type(e).__name__ != "Modeltime_too_late":
                raise
            self.default_boundary_invoked = True
            return self.default_boundary

    def values(self, t):
        """[stage, xmomentum, ymomentum] for the device at time t."""
        raise NotImplementedError


def _stage_of(value):
    try:
        return float(value)
    except TypeError:
        return float(value[0])


class Time_boundary(_TimeFunctionBoundary):
    """Dirichlet values [stage, xmomentum, ymomentum] = function(t)
    (generic_boundary_conditions.py:297-417)."""
    kind = "time"
    code = KIND_DIRICHLET

    def __init__(self, domain=None, function=None, default_boundary=None, verbose=False):
        super().__init__(domain, function, default_boundary, verbose)
        try:
            q = np.array(function(0.0), dtype=float)
        except Exception as e:
            raise Exception("Function for time boundary could not be executed:\n%s" % e)
        if q.ndim != 1 or len(q) != 3:
            raise Exception("Return value for function must be a list or an array of length 3")

    def values(self, t):
        return np.asarray(self.get_boundary_values(t), dtype=np.float64)[:3]

    def __repr__(self):
        return "Time boundary"


class Time_stage_zero_momentum_boundary(_TimeFunctionBoundary):
    """Stage = function(t), zero momentum (boundaries.py:553-636)."""
    kind = "time_stage_zero_momentum"
    code = KIND_DIRICHLET

    def values(self, t):
        return np.array([_stage_of(self.get_boundary_values(t)), 0.0, 0.0])


class Transmissive_n_momentum_zero_t_momentum_set_stage_boundary(_TimeFunctionBoundary):
    """Stage = function(t); the normal momentum of the boundary triangle's
    edge kept, the tangential set to zero (boundaries.py:384-517)."""
    kind = "transmissive_n_momentum_zero_t_momentum_set_stage"
    code = KIND_SET_STAGE_NORMAL_MOMENTUM

    def __init__(self, domain=None, function=None, default_boundary=0.0):
        super().__init__(domain, function, None)
        self.default_boundary = default_boundary

    def values(self, t):
        return np.array([_stage_of(self.get_boundary_values(t)), 0.0, 0.0])


class Transmissive_momentum_set_stage_boundary(_TimeFunctionBoundary):
    """Stage = function(t), momentum of the boundary triangle's edge
    (boundaries.py:295-381)."""
    kind = "transmissive_momentum_set_stage"
    code = KIND_SET_STAGE_TRANSMISSIVE_MOMENTUM

    def __init__(self, domain=None, function=None):
        super().__init__(domain, function, None)

    def values(self, t):
        f = self.function
        if hasattr(f, "time"):
            while t > f.time[-1]:        # roll over, as the reference does
                t -= f.time[-1]
        return np.array([_stage_of(f(t)), 0.0, 0.0])


class Transmissive_stage_zero_momentum_boundary(Boundary):
    """Stage of the boundary triangle's edge, zero momentum
    (boundaries.py:520-550)."""
    kind = "transmissive_stage_zero_momentum"
    code = KIND_TRANSMISSIVE_STAGE_ZERO_MOMENTUM

    def __init__(self, domain=None):
        if domain is None:
            raise Exception("Domain must be specified for "
                            "Transmissive_stage_zero_momentum boundary")
        self.domain = domain

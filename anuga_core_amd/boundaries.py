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

"""MI355X-native DE shallow-water step with the anuga.Domain interface.

    from anuga_core_amd import rectangular_cross_domain, Reflective_boundary
    domain = rectangular_cross_domain(100, 100, len1=100., len2=100.)
    domain.set_flow_algorithm('DE1')
    ...
    for t in domain.evolve(yieldstep=1.0, finaltime=20.0):
        print(domain.get_quantity('stage').centroid_values.max())
"""

from .boundaries import (Dirichlet_boundary, Reflective_boundary, Time_boundary,
                         Time_stage_zero_momentum_boundary, Transmissive_boundary,
                         Transmissive_momentum_set_stage_boundary,
                         Transmissive_n_momentum_zero_t_momentum_set_stage_boundary,
                         Transmissive_stage_zero_momentum_boundary)
from .mesh import Mesh, rectangular, rectangular_cross

__all__ = ["Domain", "Ensemble", "Mesh", "rectangular", "rectangular_cross",
           "rectangular_cross_domain", "rectangular_domain", "Reflective_boundary",
           "Transmissive_boundary", "Dirichlet_boundary", "Time_boundary",
           "Time_stage_zero_momentum_boundary", "Transmissive_momentum_set_stage_boundary",
           "Transmissive_n_momentum_zero_t_momentum_set_stage_boundary",
           "Transmissive_stage_zero_momentum_boundary"]


def __getattr__(name):
    # Domain needs the device library; import it on first use so that the
    # mesh and boundary modules stay usable without it.
    if name == "Domain":
        from .domain import Domain
        return Domain
    if name == "Ensemble":
        from .ensemble import Ensemble
        return Ensemble
    raise AttributeError(name)


def rectangular_cross_domain(*args, **kwargs):
    """anuga.rectangular_cross_domain (anuga/extras.py:13-47)."""
    from .domain import Domain
    kwargs.pop("verbose", None)
    dom_kw = {k: kwargs.pop(k) for k in ("precision", "reorder") if k in kwargs}
    points, vertices, boundary = rectangular_cross(*args, **kwargs)
    return Domain(points, vertices, boundary, **dom_kw)


def rectangular_domain(*args, **kwargs):
    """Domain on anuga's `rectangular` mesh (mesh_factory.py:84-135)."""
    from .domain import Domain
    dom_kw = {k: kwargs.pop(k) for k in ("precision", "reorder") if k in kwargs}
    points, vertices, boundary = rectangular(*args, **kwargs)
    return Domain(points, vertices, boundary, **dom_kw)

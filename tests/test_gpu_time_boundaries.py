"""Time-dependent and set-stage boundaries on the device (MI355X): the host
evaluates the boundary functions at each substep time (t, t + dt, t + dt/2)
and uploads them; the flux kernel applies the new kinds. Against the
oracle's runs (update_boundary at the same times): timestep sequences
identical, states bit for bit, for euler, rk2 and rk3.
"""

import math

import numpy as np
import pytest

from anuga_core_amd.boundaries import (Reflective_boundary, Time_boundary,
                                       Time_stage_zero_momentum_boundary,
                                       Transmissive_momentum_set_stage_boundary,
                                       Transmissive_n_momentum_zero_t_momentum_set_stage_boundary,
                                       Transmissive_stage_zero_momentum_boundary)
from anuga_core_amd.domain import Domain
from anuga_core_amd.mesh import Mesh, rectangular_cross
from oracle.kernels import OracleKernels
from oracle.oracle_domain import OracleDomain

pytestmark = pytest.mark.gpu


def _setup(dom, alg, variant):
    mesh = dom.mesh
    x, y = mesh.centroid_coordinates.T
    z = -0.02 * x + 0.05 * np.sin(y / 7.0)
    q = dict(elevation=z, stage=np.maximum(0.4 + 0.0 * x, z), friction=np.full(len(x), 0.02))
    dom.set_flow_algorithm(alg)
    for name, v in q.items():
        if hasattr(dom, "set_centroid_values"):
            dom.set_centroid_values(name, v)
        else:
            dom.set_quantity(name, v, location="centroids")
    if variant == 0:
        bc = {"left": Time_boundary(dom, lambda t: [0.5 + 0.2 * math.sin(t / 3.0), 0.05 * t, 0.0]),
              "right": Transmissive_n_momentum_zero_t_momentum_set_stage_boundary(
                  dom, lambda t: 0.35 + 0.01 * t),
              "top": Transmissive_stage_zero_momentum_boundary(dom),
              "bottom": Reflective_boundary(dom)}
    else:
        bc = {"left": Time_stage_zero_momentum_boundary(dom, lambda t: [0.6 - 0.01 * t]),
              "right": Transmissive_momentum_set_stage_boundary(dom, lambda t: 0.3),
              "top": Reflective_boundary(dom), "bottom": Reflective_boundary(dom)}
    dom.set_boundary(bc)
    return dom


@pytest.mark.parametrize("alg", ["DE0", "DE1", "DE2"])
@pytest.mark.parametrize("variant", [0, 1])
def test_time_boundaries_bitwise_vs_oracle(alg, variant):
    print("case time boundaries %s variant %d" % (alg, variant), flush=True)
    p, v, b = rectangular_cross(24, 12, 60.0, 30.0)
    mesh = Mesh(p, v, b)
    d = _setup(Domain(mesh=mesh), alg, variant)
    o = _setup(OracleDomain(mesh, OracleKernels(gather=True)), alg, variant)
    for _ in d.evolve(yieldstep=2.0, finaltime=8.0):
        pass
    for _ in o.evolve(yieldstep=2.0, finaltime=8.0):
        pass
    assert d.dt_history == o.dt_history
    for name, oname in (("stage", "stage_c"), ("xmomentum", "xmom_c"), ("ymomentum", "ymom_c")):
        assert np.array_equal(d.get_quantity(name).centroid_values, getattr(o, oname)), name
    # water came in through the time boundary
    assert d.get_water_volume() != pytest.approx(0.0)

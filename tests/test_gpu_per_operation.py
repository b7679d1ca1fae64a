"""The per-operation entry points on the device (MI355X), driven in the exact
call order of the reference's evolve_one_euler_step / evolve_one_rk2_step /
evolve_one_rk3_step (generic_domain.py:1914-2179): backup, distribute,
update_boundary, compute_fluxes_ext_central, compute_forcing_terms,
update_timestep (host), Quantity.update x3 + fix_negative_cells, saxpy.

Against the oracle's runs of the same scenarios (the reference's
orchestration over the C restatement): identical timestep sequences and
states bit for bit; against the fused steps of the same library: the same
bits; and the reference's call-count substep rule for flux_timestep.
"""

import ctypes as C
import math

import numpy as np
import pytest

from anuga_core_amd import _native
from anuga_core_amd.boundaries import Reflective_boundary, Time_boundary
from anuga_core_amd.domain import Domain
from anuga_core_amd.mesh import Mesh, rectangular_cross
from oracle.kernels import OracleKernels
from oracle.oracle_domain import OracleDomain
from tests import scenarios as S

pytestmark = pytest.mark.gpu

NAMES = (("stage", "stage_c"), ("xmomentum", "xmom_c"), ("ymomentum", "ymom_c"))


def _run(dom, yieldstep, finaltime):
    for _ in dom.evolve(yieldstep=yieldstep, finaltime=finaltime):
        pass


@pytest.mark.parametrize("case", [
    ("dam_break_de0", dict(m=30, n=30), "DE0", 1.0, 3.0),
    ("hills_de1", dict(m=24, n=24), "DE1", 5.0, 30.0),
    ("hills_de1", dict(m=20, n=20), "DE2", 5.0, 20.0),
])
def test_per_operation_steps_bitwise_vs_oracle_and_fused(case):
    builder, kw, alg, ys, ft = case
    print("case per-operation %s %s" % (builder, alg), flush=True)
    mesh, _, q, bc = getattr(S, builder)(**kw)
    o = S.setup(OracleDomain(mesh, OracleKernels(gather=True)), alg, q, bc)
    d = S.setup(Domain(mesh=mesh), alg, q, bc)
    d.set_step_mode("per_operation")
    f = S.setup(Domain(mesh=mesh), alg, q, bc)
    _run(o, ys, ft)
    _run(d, ys, ft)
    _run(f, ys, ft)
    print("steps", len(d.dt_history), flush=True)
    assert len(d.dt_history) > 10
    assert d.dt_history == o.dt_history, "timestep sequence differs from the oracle"
    assert d.dt_history == f.dt_history, "timestep sequence differs from the fused steps"
    for name, oname in NAMES:
        a = d.get_quantity(name).centroid_values
        assert np.array_equal(a, getattr(o, oname)), name
        assert np.array_equal(a, f.get_quantity(name).centroid_values), name
    bo, bd = o.boundary_flux_integral, d.boundary_flux_integral
    assert abs(bd - bo) <= 1e-12 * max(abs(bo), 1.0)
    assert d.boundary_flux_integral == f.boundary_flux_integral


def test_per_operation_time_boundary_rk2_vs_oracle():
    """A time-dependent boundary: update_boundary uploads the values of the
    current time before each compute_fluxes, as the reference evaluates them."""
    print("case per-operation time boundary", flush=True)
    p, v, b = rectangular_cross(20, 10, 60.0, 30.0)
    mesh = Mesh(p, v, b)

    def setup(dom):
        x, y = mesh.centroid_coordinates.T
        z = -0.02 * x + 0.05 * np.sin(y / 7.0)
        dom.set_flow_algorithm("DE1")
        for name, val in (("elevation", z), ("stage", np.maximum(0.4 + 0.0 * x, z)),
                          ("friction", np.full(len(x), 0.02))):
            if hasattr(dom, "set_centroid_values"):
                dom.set_centroid_values(name, val)
            else:
                dom.set_quantity(name, val, location="centroids")
        dom.set_boundary({"left": Time_boundary(dom, lambda t: [0.5 + 0.2 * math.sin(t / 3.0),
                                                                0.05 * t, 0.0]),
                          "right": Reflective_boundary(dom), "top": Reflective_boundary(dom),
                          "bottom": Reflective_boundary(dom)})
        return dom

    d = setup(Domain(mesh=mesh))
    d.set_step_mode("per_operation")
    o = setup(OracleDomain(mesh, OracleKernels(gather=True)))
    _run(d, 2.0, 6.0)
    _run(o, 2.0, 6.0)
    assert d.dt_history == o.dt_history
    for name, oname in NAMES:
        assert np.array_equal(d.get_quantity(name).centroid_values, getattr(o, oname)), name


def test_flux_timestep_substep_rule_and_errors():
    """compute_fluxes_ext_central returns the CFL timestep on the first call of
    a step and the timestep passed in on the others (sw_domain_orig.c:1146-1149);
    a saxpy without a backup and a flux call on stale edge values are refused."""
    print("case per-operation substep rule", flush=True)
    mesh, _, q, bc = S.hills_de1(m=12, n=12)
    d = S.setup(Domain(mesh=mesh), "DE1", q, bc)
    L, h = _native.lib(), d._handle
    with pytest.raises(_native.DeviceError):
        d.saxpy_conserved_quantities(0.5, 0.5)
    d.timestep = 0.0
    d.update_conserved_quantities()        # makes the edge values stale
    ft = C.c_double()
    assert L.anuga_compute_fluxes_ext_central(h, C.c_double(7.0), C.byref(ft)) == \
        _native.ANUGA_E_STATE
    d.distribute_to_vertices_and_edges()
    assert L.anuga_compute_fluxes_ext_central(h, C.c_double(7.0), C.byref(ft)) == 0
    first = ft.value
    assert 0.0 < first < 7.0
    assert L.anuga_compute_fluxes_ext_central(h, C.c_double(7.0), C.byref(ft)) == 0
    assert ft.value == 7.0                 # the second substep of rk2
    assert L.anuga_compute_fluxes_ext_central(h, C.c_double(7.0), C.byref(ft)) == 0
    assert ft.value == first               # a new step

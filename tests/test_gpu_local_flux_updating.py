"""Local extrapolation and flux updating on the device (MI355X) against the
oracle's restatement, which equals the reference's compiled C bit for bit
(tests/test_local_flux_updating.py).

* Identical dt sequences and states bit for bit, DE0 with 2**3 and 2**8
  levels, structured and unstructured meshes, fused and per-operation step
  modes; the boundary flux integral to rounding (its sum order differs).
* A 40,000-triangle dam break over many blocks and frequency cycles.
* The reference's acceptance test (test_local_extrapolation_and_flux_updating.py:61-88)
  on the device.
* max_flux_update_frequency set directly, as the reference's C reads it.
"""

import numpy as np
import pytest

from anuga_core_amd.domain import Domain
from oracle.kernels import OracleKernels
from oracle.oracle_domain import OracleDomain
from tests import scenarios as S

pytestmark = pytest.mark.gpu


def _pair(scenario, nlevels, step_mode="fused"):
    mesh, alg, q, bc = scenario
    o = S.setup(OracleDomain(mesh, OracleKernels()), alg, q, bc)
    d = S.setup(Domain(mesh=mesh), alg, q, bc)
    d.set_step_mode(step_mode)
    if nlevels is not None:
        o.set_local_extrapolation_and_flux_updating(nlevels=nlevels)
        d.set_local_extrapolation_and_flux_updating(nlevels=nlevels)
    return o, d


def _compare(o, d, yieldstep, finaltime):
    for _ in d.evolve(yieldstep=yieldstep, finaltime=finaltime):
        pass
    for _ in o.evolve(yieldstep=yieldstep, finaltime=finaltime):
        pass
    assert d.dt_history == o.dt_history
    for name, oname in (("stage", "stage_c"), ("xmomentum", "xmom_c"), ("ymomentum", "ymom_c")):
        assert np.array_equal(d.get_quantity(name).centroid_values, getattr(o, oname)), name
    assert d.get_boundary_flux_integral() == pytest.approx(o.boundary_flux_integral,
                                                           rel=1e-12, abs=1e-12)


@pytest.mark.parametrize("step_mode", ["fused", "per_operation"])
@pytest.mark.parametrize("graded", [False, True])
@pytest.mark.parametrize("nlevels", [3, 8])
def test_device_matches_oracle(nlevels, graded, step_mode):
    print("case local flux updating nlevels %d graded %s %s" % (nlevels, graded, step_mode),
          flush=True)
    o, d = _pair(S.inflow_over_slope(24, 24, graded=graded, fine_area=8.0), nlevels, step_mode)
    assert not d._device_stepping()
    _compare(o, d, 2.0, 12.0)
    assert len(d.dt_history) > 40


def test_large_dam_break_vs_oracle():
    print("case local flux updating dam break 40k", flush=True)
    o, d = _pair(S.dam_break_de0(m=100, n=100), 4)
    _compare(o, d, 1.0, 3.0)
    assert len(d.dt_history) > 100


def test_reference_acceptance_de0():
    print("case local flux updating acceptance", flush=True)
    runs = []
    for nlevels in (None, 8):
        mesh, alg, q, bc = S.inflow_over_slope(graded=True)
        d = S.setup(Domain(mesh=mesh), alg, q, bc)
        if nlevels is not None:
            d.set_local_extrapolation_and_flux_updating(nlevels=nlevels)
        for _ in d.evolve(yieldstep=0.1, finaltime=20.0):
            pass
        runs.append((d.get_water_volume(), d.get_boundary_flux_integral(),
                     d.get_quantity("stage").centroid_values.copy()))
        d.close()
    (v1, b1, s1), (v2, b2, s2) = runs
    assert np.allclose(v1, v2, rtol=0.05)
    assert np.allclose(v1, b1)
    assert np.allclose(v2, b2)
    assert np.all(np.abs(s1 - s2) < 0.02)
    assert not np.array_equal(s1, s2)


def test_attribute_set_directly():
    print("case local flux updating attribute", flush=True)
    o, d = _pair(S.inflow_over_slope(24, 24, fine_area=8.0), None)
    o.set_local_extrapolation_and_flux_updating(nlevels=2)
    d.max_flux_update_frequency = 4
    _compare(o, d, 2.0, 6.0)

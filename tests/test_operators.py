"""Rate_operator (anuga/operators/rate_operators.py:149-260) on the oracle
(CPU): closed forms of rain on still water (every triangle rises by
factor * dt * rate each step; the added volume is the rate times area times
time), of extraction (the loss stops at the bed, the momenta scale with the
remaining depth), the reference's time of evaluation inside euler / rk2
steps, and the host-side region selection."""

import numpy as np
import pytest

from anuga_core_amd.mesh import Mesh, rectangular_cross
from anuga_core_amd.operators import _inside_polygon
from oracle.kernels import OracleKernels
from oracle.oracle_domain import OracleDomain, OracleRateOperator
from tests import scenarios as S


def _basin(alg="DE0", depth=1.0, m=6):
    p, v, b = rectangular_cross(m, m, 30.0, 30.0)
    mesh = Mesh(p, v, b)
    N = mesh.number_of_triangles
    q = dict(elevation=np.zeros(N), stage=np.full(N, depth), friction=np.zeros(N))
    o = S.setup(OracleDomain(mesh, OracleKernels()), alg, q,
                {t: "reflective" for t in ("left", "right", "top", "bottom")})
    return o


@pytest.mark.parametrize("alg", ["DE0", "DE1"])
def test_uniform_rain_on_still_water(alg):
    o = _basin(alg)
    op = OracleRateOperator(rate=2.0e-3, factor=0.5)
    o.operators.append(op)
    v0 = o.areas.sum() * 1.0
    expect = 1.0
    for _ in o.evolve(yieldstep=5.0, finaltime=10.0):
        pass
    for dt in o.dt_history:
        expect = expect + 0.5 * dt * 2.0e-3
    assert np.allclose(o.stage_c, expect, rtol=0, atol=1e-12)
    area = o.areas.sum()
    assert o.get_water_volume() - v0 == pytest.approx(0.5 * 2.0e-3 * 10.0 * area, rel=1e-9)
    assert op.cumulative_influx == pytest.approx(0.5 * 2.0e-3 * 10.0 * area, rel=1e-9)


def test_extraction_stops_at_the_bed():
    o = _basin("DE0", depth=0.01)
    o.xmom_c[:] = 0.0
    idx = np.arange(0, o.N, 3)
    op = OracleRateOperator(indices=idx, rate=-1.0e-3)
    o.operators.append(op)
    for _ in o.evolve(yieldstep=10.0, finaltime=30.0):
        pass
    h = o.stage_c - o.elev_c
    assert h.min() >= -1e-15
    assert op.cumulative_influx < 0.0
    # the loss never exceeds what was there: total volume stays >= 0
    assert o.get_water_volume() >= 0.0


def test_rate_evaluated_at_the_reference_times():
    for alg, after in (("DE0", 0.0), ("DE1", 1.0)):
        o = _basin(alg)
        times = []

        def rate(t):
            times.append(t)
            return 1e-4
        o.operators.append(OracleRateOperator(rate=rate))
        for _ in o.evolve(yieldstep=2.0, finaltime=4.0):
            pass
        starts = np.concatenate([[0.0], np.cumsum(o.dt_history)[:-1]])
        expect = starts + after * np.asarray(o.dt_history)
        assert np.allclose(times, expect, rtol=0, atol=1e-13), alg


def test_inside_polygon():
    pts = np.array([[0.5, 0.5], [1.5, 0.5], [1.0, 1.0], [0.0, 0.25], [2.0, 2.0]])
    square = [[0, 0], [1, 0], [1, 1], [0, 1]]
    assert _inside_polygon(pts, square).tolist() == [True, False, True, True, False]
    tri = [[0, 0], [2, 0], [0, 2]]
    assert _inside_polygon(pts, tri).tolist() == [True, True, True, True, False]


def test_set_stage_holds_its_region():
    """Set_stage_operator on the oracle: after every step the region's stage
    is the set value (the operator runs after the step), the rest of the
    basin fills from it, momenta in the region are left as the step made
    them; a function of t is taken at the domain's time after an rk2 step,
    which for the last step is the final time."""
    from oracle.oracle_domain import OracleSetStageOperator
    o = _basin("DE1", depth=1.0, m=8)
    p, v, b = rectangular_cross(8, 8, 30.0, 30.0)
    mesh = Mesh(p, v, b)
    region = np.nonzero(_inside_polygon(mesh.centroid_coordinates,
                                        [[0, 0], [10, 0], [10, 10], [0, 10]]))[0]
    assert len(region) > 0
    o.operators.append(OracleSetStageOperator(region, 1.5))
    for _ in o.evolve(yieldstep=1.0, finaltime=3.0):
        pass
    assert np.all(o.stage_c[region] == 1.5)
    outside = np.setdiff1d(np.arange(mesh.number_of_triangles), region)
    assert o.stage_c[outside].max() > 1.0 + 1e-6          # water spreads from the region
    assert np.any(o.xmom_c[region] != 0.0)                 # momenta not reset
    o2 = _basin("DE1", depth=1.0, m=8)
    f = lambda t: 1.0 + 0.1 * t
    o2.operators.append(OracleSetStageOperator(region, f))
    for _ in o2.evolve(yieldstep=1.0, finaltime=2.0):
        pass
    assert np.all(o2.stage_c[region] == f(o2.get_time()))
    assert o2.get_time() == 2.0
    # f(x, y): evaluated at the region's centroids
    o3 = _basin("DE0", depth=1.0, m=8)
    g = lambda x, y: 1.0 + 0.01 * x
    o3.operators.append(OracleSetStageOperator(region, g, mesh.centroid_coordinates))
    for _ in o3.evolve(yieldstep=0.5, finaltime=1.0):
        pass
    assert np.array_equal(o3.stage_c[region], g(*mesh.centroid_coordinates[region].T))

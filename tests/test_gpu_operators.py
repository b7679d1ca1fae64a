"""Rate operators on the device (MI355X) against the oracle's restatement of
Rate_operator.__call__ (tests/test_operators.py): static rates applied inside
the captured step (device-controlled runs), time-dependent rates evaluated on
the host at the reference's times (host-sequenced runs), negative rates.
Timestep sequences identical, states bit for bit.
"""

import math

import numpy as np
import pytest

from anuga_core_amd.domain import Domain
from anuga_core_amd.operators import Rate_operator
from oracle.kernels import OracleKernels
from oracle.oracle_domain import OracleDomain, OracleRateOperator
from tests import scenarios as S

pytestmark = pytest.mark.gpu


def _pair(alg):
    mesh, _, q, bc = S.dam_break_de0(m=16, n=12)
    o = S.setup(OracleDomain(mesh, OracleKernels(gather=True)), alg, q, bc)
    d = S.setup(Domain(mesh=mesh), alg, q, bc)
    return mesh, o, d


def _run_and_compare(o, d, ys, ft):
    for _ in d.evolve(yieldstep=ys, finaltime=ft):
        pass
    for _ in o.evolve(yieldstep=ys, finaltime=ft):
        pass
    assert d.dt_history == o.dt_history
    for name, oname in (("stage", "stage_c"), ("xmomentum", "xmom_c"), ("ymomentum", "ymom_c")):
        assert np.array_equal(d.get_quantity(name).centroid_values, getattr(o, oname)), name


@pytest.mark.parametrize("alg", ["DE0", "DE1", "DE2"])
def test_static_rain_and_extraction_vs_oracle(alg):
    print("case static rate operators %s" % alg, flush=True)
    mesh, o, d = _pair(alg)
    cen = mesh.centroid_coordinates
    rain = Rate_operator(d, rate=5.0e-3, factor=0.8,
                         polygon=[[10, 10], [60, 10], [60, 60], [10, 60]])
    well = Rate_operator(d, rate=-2.0e-2, center=[25.0, 50.0], radius=12.0)
    spatial = Rate_operator(d, rate=lambda x, y: 1e-4 * (x / 100.0))
    o.operators += [OracleRateOperator(rain.indices, 5.0e-3, 0.8),
                    OracleRateOperator(well.indices, -2.0e-2),
                    OracleRateOperator(None, 1e-4 * (cen[:, 0] / 100.0))]
    assert len(rain.indices) > 0 and len(well.indices) > 0
    assert d._device_stepping()
    _run_and_compare(o, d, 1.0, 4.0)
    infl = [op.get_cumulative_influx() for op in (rain, well, spatial)]
    for a, b in zip(infl, o.operators):
        assert a == pytest.approx(b.cumulative_influx, rel=1e-12, abs=1e-15)


@pytest.mark.parametrize("alg", ["DE0", "DE1"])
def test_time_dependent_rate_vs_oracle(alg):
    print("case time-dependent rate operator %s" % alg, flush=True)
    mesh, o, d = _pair(alg)
    f = lambda t: 2.0e-3 * (1.0 + math.sin(t))
    Rate_operator(d, rate=f, indices=np.arange(0, mesh.number_of_triangles, 2))
    o.operators.append(OracleRateOperator(np.arange(0, mesh.number_of_triangles, 2), f))
    assert not d._device_stepping()
    _run_and_compare(o, d, 1.0, 3.0)


@pytest.mark.parametrize("alg", ["DE0", "DE1"])
def test_inlets_vs_oracle(alg):
    """Inlet_operator (levelled on the device) next to a static rain
    operator, against the oracle's numpy restatement."""
    print("case inlets %s" % alg, flush=True)
    from anuga_core_amd.operators import Inlet_operator, _line_triangles
    from oracle.oracle_domain import OracleInletOperator
    mesh, o, d = _pair(alg)
    l1 = _line_triangles(mesh, [[30.0, 20.0], [40.0, 20.0]])
    l2 = _line_triangles(mesh, [[70.0, 60.0], [80.0, 60.0]])
    Inlet_operator(d, [[30.0, 20.0], [40.0, 20.0]], Q=2.0)
    Rate_operator(d, rate=1e-3, polygon=[[0, 0], [50, 0], [50, 50], [0, 50]])
    Inlet_operator(d, l2, Q=lambda t: -0.5 + 0.2 * t, velocity=[0.1, 0.0])
    rain = d.fractional_step_operators[1]
    o.operators += [OracleInletOperator(l1, 2.0), OracleRateOperator(rain.indices, 1e-3),
                    OracleInletOperator(l2, lambda t: -0.5 + 0.2 * t, velocity=[0.1, 0.0])]
    assert not d._device_stepping()
    _run_and_compare(o, d, 1.0, 4.0)


@pytest.mark.parametrize("alg", ["DE0", "DE1"])
def test_set_stage_vs_oracle(alg):
    """Set_stage_operator: a constant over a polygon and a function of (x, y)
    over a circle applied inside the captured step, next to rain; then a
    function of t (host-sequenced). Bit for bit against the oracle."""
    print("case set stage %s" % alg, flush=True)
    from anuga_core_amd.operators import Set_stage_operator
    from oracle.oracle_domain import OracleSetStageOperator
    mesh, o, d = _pair(alg)
    cen = mesh.centroid_coordinates
    pond = Set_stage_operator(d, stage=0.4, polygon=[[60, 20], [90, 20], [90, 50], [60, 50]])
    g = lambda x, y: 0.2 + 0.001 * y
    hill = Set_stage_operator(d, stage=g, center=[30.0, 80.0], radius=10.0)
    rain = Rate_operator(d, rate=1.0e-3)
    o.operators += [OracleSetStageOperator(pond.indices, 0.4),
                    OracleSetStageOperator(hill.indices, g, cen),
                    OracleRateOperator(None, 1.0e-3)]
    assert len(pond.indices) > 0 and len(hill.indices) > 0
    assert d._device_stepping()
    _run_and_compare(o, d, 1.0, 3.0)

    mesh, o, d = _pair(alg)
    f = lambda t: 0.3 + 0.05 * math.sin(t)
    tide = Set_stage_operator(d, stage=f, indices=np.arange(0, 40))
    o.operators.append(OracleSetStageOperator(np.arange(0, 40), f))
    assert not d._device_stepping()
    _run_and_compare(o, d, 1.0, 3.0)
    if alg == "DE1":   # rk2 applies operators at the step's end time
        assert np.all(d.get_quantity("stage").centroid_values[tide.indices] == f(d.get_time()))


@pytest.mark.parametrize("alg", ["DE0", "DE1"])
def test_device_inlets_match_host_inlets(alg):
    """The device's inlet levelling (one workgroup per inlet, after every
    step) against the same operators applied on the host to the region's
    state read back (the reference's numpy arithmetic): the same bits on
    this equal-area mesh, for a filling, a draining and a time-dependent
    inlet with each velocity mode; a constant-Q domain keeps runs of steps
    on the device."""
    print("case device inlets %s" % alg, flush=True)
    from anuga_core_amd.operators import Inlet_operator, _line_triangles
    runs = []
    for device in (True, False):
        mesh, _, q, bc = S.dam_break_de0(m=16, n=12)
        d = S.setup(Domain(mesh=mesh), alg, q, bc)
        a = Inlet_operator(d, [[30.0, 20.0], [40.0, 20.0]], Q=2.0, device=device)
        b = Inlet_operator(d, _line_triangles(mesh, [[10.0, 10.0], [25.0, 12.0]]), Q=-0.3,
                           zero_velocity=True, device=device)
        c = Inlet_operator(d, _line_triangles(mesh, [[45.0, 5.0], [45.0, 15.0]]),
                           Q=lambda t: 0.5 + 0.2 * t, velocity=[0.1, -0.05], device=device)
        if device:
            assert not a.host_side and not a.time_dependent and c.time_dependent
        for _ in d.evolve(yieldstep=1.0, finaltime=4.0):
            pass
        runs.append((list(d.dt_history), [d.get_quantity(n).centroid_values.copy() for n in
                                          ("stage", "xmomentum", "ymomentum")],
                     [op.get_cumulative_influx() for op in (a, b, c)]))
        d.close()
    assert runs[0][0] == runs[1][0]
    for x, y in zip(runs[0][1], runs[1][1]):
        assert np.array_equal(x, y)
    assert runs[0][2] == runs[1][2]


def test_constant_inlet_runs_on_the_device():
    print("case device inlet device-stepping", flush=True)
    from anuga_core_amd.operators import Inlet_operator
    mesh, _, q, bc = S.dam_break_de0(m=16, n=12)
    d = S.setup(Domain(mesh=mesh), "DE1", q, bc)
    v0 = d.get_water_volume()
    op = Inlet_operator(d, [[30.0, 20.0], [40.0, 20.0]], Q=2.0)
    assert d._device_stepping()
    d.evolve_steps(40)
    added = op.get_cumulative_influx()
    assert added == pytest.approx(2.0 * d.get_time(), rel=1e-12)
    assert d.get_water_volume() - v0 == pytest.approx(added, rel=1e-9)

"""Inlet_operator on the oracle (CPU): the reference's known answers of
anuga/structures/tests/test_inlet_operator.py (the volume added equals
Q * t for constant and time-varying Q, a negative Q drains at most what is
there) on its set-up, and the level surface that set_stages_evenly builds."""

import numpy as np
import pytest

from anuga_core_amd.mesh import Mesh, rectangular_cross
from anuga_core_amd.operators import _line_triangles
from oracle.kernels import OracleKernels
from oracle.oracle_domain import OracleDomain, OracleInletOperator
from tests import scenarios as S


def _domain(stage_0=11.0, stage_1=10.0, elevation_0=10.0, elevation_1=10.0, alg="DE0"):
    """_create_domain of test_inlet_operator.py:37-89: 200 m square, 10 m cells."""
    mesh = Mesh(*rectangular_cross(20, 20, len1=200.0, len2=200.0))
    x = mesh.centroid_coordinates[:, 0]
    z = np.where(x > 100.0, elevation_1, elevation_0)
    w = np.where(x > 100.0, stage_1, stage_0)
    N = mesh.number_of_triangles
    o = S.setup(OracleDomain(mesh, OracleKernels()), alg,
                dict(elevation=z, stage=w, friction=np.zeros(N)),
                {t: "reflective" for t in ("left", "right", "top", "bottom")})
    return mesh, o


def _volume(o):
    return float(np.sum((o.stage_c - o.elev_c) * o.areas))


@pytest.mark.parametrize("alg", ["DE0", "DE1"])
def test_inlet_constant_Q(alg):
    mesh, o = _domain(alg=alg)
    v0 = _volume(o)
    i1 = OracleInletOperator(_line_triangles(mesh, [[95.0, 10.0], [105.0, 10.0]]), 5.0)
    i2 = OracleInletOperator(_line_triangles(mesh, [[10.0, 90.0], [20.0, 90.0]]), 10.0)
    o.operators += [i1, i2]
    for _ in o.evolve(yieldstep=1.0, finaltime=3.0):
        pass
    assert _volume(o) - v0 == pytest.approx(15.0 * 3.0, rel=1e-8)
    assert i1.volume_integral + i2.volume_integral == pytest.approx(45.0, rel=1e-8)


def test_inlet_variable_Q_and_negative_Q():
    mesh, o = _domain()
    v0 = _volume(o)
    line = _line_triangles(mesh, [[95.0, 10.0], [105.0, 10.0]])
    o.operators.append(OracleInletOperator(line, lambda t: 5.0 + 2.0 * t))
    for _ in o.evolve(yieldstep=1.0, finaltime=3.0):
        pass
    # integral of 5 + 2t over [0, 3], exact for the trapezoidal Q of each step
    assert _volume(o) - v0 == pytest.approx(15.0 + 9.0, rel=1e-8)
    # draining more than the inlet holds empties it and no more
    mesh, o = _domain(stage_0=10.01, stage_1=10.01)
    drain = OracleInletOperator(line, -1.0e3)
    o.operators.append(drain)
    for _ in o.evolve(yieldstep=1.0, finaltime=1.0):
        pass
    h = (o.stage_c - o.elev_c)[line]
    assert np.all(h >= 0.0)
    assert drain.volume_integral < 0.0


def test_stages_evenly_make_a_level_surface():
    mesh, o = _domain(stage_0=10.0, stage_1=10.0)
    line = _line_triangles(mesh, [[95.0, 10.0], [105.0, 10.0]])
    o.stage_c[line] += np.linspace(0.0, 0.3, len(line))      # uneven water
    o.timestep = 1.0
    op = OracleInletOperator(line, 2.0)
    before = (o.stage_c - o.elev_c)[line] @ o.areas[line]
    op.apply(o)
    w = o.stage_c[line]
    assert (o.stage_c - o.elev_c)[line] @ o.areas[line] == pytest.approx(before + 2.0, rel=1e-12)
    # the filled cells share one stage; the rest stay above it
    top = w.min()
    assert np.all(np.isclose(w, top, rtol=0, atol=1e-12) | (w > top))

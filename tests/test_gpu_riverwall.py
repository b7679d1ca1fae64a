"""Riverwalls on the device (MI355X): anuga_set_riverwalls and the riverwall
instantiation of the FP64 flux kernel against the oracle's edge-once loop
(itself equal to the reference's compiled C bit for bit, test_riverwall.py),
with the correctly rounded pow on both sides: timestep sequences identical,
centroid states bit for bit. Also the reference's riverwall tests
(test_riverwall_structure.py:152-500) on the device, and the paths that
refuse walls.
"""

import numpy as np
import pytest

from anuga_core_amd.domain import Domain
from oracle.kernels import OracleKernels
from oracle.oracle_domain import G, OracleDomain
from tests import scenarios as S
from tests.test_riverwall import RUNS

pytestmark = pytest.mark.gpu


def _pair(builder, kw, flags=None, reorder=True):
    mesh, alg, q, bc, walls, par = builder(**kw)
    o = S.setup(OracleDomain(mesh, OracleKernels(gather=False)), alg, q, bc)
    d = S.setup(Domain(mesh=mesh, reorder=reorder), alg, q, bc)
    eft, elev, rows, hp = S.riverwall_arrays(mesh, walls, par)
    if flags is not None:
        eft, elev, rows = flags(mesh, eft, elev, rows)
    o.set_riverwalls(eft, elev, rows, hp)
    if flags is None:
        d.create_riverwalls(walls, par, verbose=False)
        assert np.array_equal(d.edge_flux_type, eft)
    else:
        d._set_riverwalls(eft, elev, rows, hp)
    return o, d


@pytest.mark.parametrize("builder,kw,ys,ft,flags", RUNS)
def test_riverwall_runs_bitwise_vs_oracle(builder, kw, ys, ft, flags):
    print("case riverwall %s %s one_sided=%s" % (builder.__name__, kw, flags is not None),
          flush=True)
    o, d = _pair(builder, kw, flags)
    for _ in d.evolve(yieldstep=ys, finaltime=ft):
        pass
    for _ in o.evolve(yieldstep=ys, finaltime=ft):
        pass
    assert d.dt_history == o.dt_history
    for name, oname in (("stage", "stage_c"), ("xmomentum", "xmom_c"), ("ymomentum", "ymom_c")):
        assert np.array_equal(d.get_quantity(name).centroid_values, getattr(o, oname)), name


def test_riverwall_host_sequenced_caller_order():
    """Host-sequenced substeps (one library call each) in the caller's own
    triangle order give the same run as the oracle."""
    print("case riverwall host-sequenced, no reorder", flush=True)
    o, d = _pair(S.riverwall_levees, dict(alg="DE1"), reorder=False)
    d.device_timestep = False
    for _ in d.evolve(yieldstep=2.0, finaltime=6.0):
        pass
    for _ in o.evolve(yieldstep=2.0, finaltime=6.0):
        pass
    assert d.dt_history == o.dt_history
    assert np.array_equal(d.get_quantity("stage").centroid_values, o.stage_c)
    assert np.array_equal(d.get_quantity("xmomentum").centroid_values, o.xmom_c)


@pytest.mark.parametrize("alg", ["DE0", "DE1"])
def test_reference_riverwall_physics_on_device(alg):
    print("case riverwall physics %s" % alg, flush=True)
    L, ft = 100.0, 1.0e-3

    def run(**kw):
        mesh, a, q, bc, walls, par = S.riverwall_structure(alg=alg, **kw)
        d = S.setup(Domain(mesh=mesh), a, q, bc)
        d.create_riverwalls(walls, par, verbose=False)
        return d, q, mesh

    def land(d, mesh):
        x = mesh.centroid_coordinates[:, 0]
        h = d.get_quantity("stage").centroid_values - d.get_quantity("elevation").centroid_values
        return float(np.sum((h * (h > 0) * mesh.areas)[x < 50.0]))

    d, q, mesh = run(wall_height=-0.2, ocean=-0.3)
    for _ in d.evolve(yieldstep=10.0, finaltime=10.0):
        pass
    assert land(d, mesh) == pytest.approx(0.0, abs=1e-12)
    for qf in (1.0, 2.0):
        d, q, mesh = run(wall_height=2.0, ocean=2.5, qfactor=qf)
        for _ in d.evolve(yieldstep=1e-4, finaltime=ft):
            pass
        theory = qf * ft * L * 2.0 / 3.0 * 0.5 * (2.0 / 3.0 * G * 0.5) ** 0.5
        assert land(d, mesh) == pytest.approx(theory, rel=1e-3)
    d, q, mesh = run(wall_height=20.0, ocean=20.5, land=20.44)
    v0 = land(d, mesh)
    for _ in d.evolve(yieldstep=1e-4, finaltime=ft):
        pass
    Q1 = 2.0 / 3.0 * 0.5 * (2.0 / 3.0 * G * 0.5) ** 0.5
    Q2 = 2.0 / 3.0 * 0.44 * (2.0 / 3.0 * G * 0.44) ** 0.5
    assert land(d, mesh) - v0 == pytest.approx(ft * L * Q1 * (1 - Q2 / Q1) ** 0.385, rel=1e-3)


def test_riverwalls_refused_where_unsupported():
    print("case riverwall refusals", flush=True)
    mesh, alg, q, bc, walls, par = S.riverwall_structure()
    d = S.setup(Domain(mesh=mesh, precision="mixed"), alg, q, bc)
    with pytest.raises(Exception, match="fp64"):
        d.create_riverwalls(walls, par, verbose=False)
    d.close()
    d = S.setup(Domain(mesh=mesh), "DE1", q, bc)
    d.create_riverwalls(walls, par, verbose=False)
    with pytest.raises(Exception, match="riverwalls"):
        d.adjoint_record(5, 0.01, [0, 1], 1)
    # removing the walls restores the plain kernel's results
    d.riverwallData.__init__(d)
    d._set_riverwalls(np.zeros(3 * mesh.number_of_triangles, np.int64), [], [],
                      [[1.0, 0.9, 0.95, 1.0, 1.5]])
    o = S.setup(OracleDomain(mesh, OracleKernels()), "DE1", q, bc)
    for _ in d.evolve(yieldstep=0.5, finaltime=1.0):
        pass
    for _ in o.evolve(yieldstep=0.5, finaltime=1.0):
        pass
    assert np.array_equal(d.get_quantity("stage").centroid_values, o.stage_c)

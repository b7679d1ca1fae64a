"""The unstructured graded mesh (anuga_core_amd.mesh.graded_square), CPU:
valid connectivity and geometry, and on it the per-triangle gather flux
equals the reference's compute-once-and-mirror loop bit for bit (both in the
oracle), the property the device kernels rely on."""

import numpy as np

from anuga_core_amd.mesh import Mesh, graded_square
from oracle.kernels import OracleKernels
from oracle.oracle_domain import OracleDomain
from tests import scenarios as S


def test_graded_square_mesh_is_valid():
    p, tri, b = graded_square(1000.0, 40.0, 300.0, (300, 300, 700, 700), seed=3)
    m = Mesh(p, tri, b)
    assert np.all(m.areas > 0)
    assert abs(m.areas.sum() - 1000.0 * 1000.0) < 1e-6 * 1e6          # covers the square
    fine = (m.centroid_coordinates >= 320).all(1) & (m.centroid_coordinates <= 680).all(1)
    assert 30 < np.median(m.areas[fine]) < 50 and 200 < np.median(m.areas[~fine]) < 400
    # neighbours are mutual and boundary edges all tagged by side
    nb, ne = m.neighbours, m.neighbour_edges
    k = np.repeat(np.arange(len(nb)), 3)[nb.ravel() >= 0]
    n, e = nb.ravel()[nb.ravel() >= 0], ne.ravel()[nb.ravel() >= 0]
    assert np.array_equal(nb[n, e], k)
    assert set(b.values()) == {"left", "right", "bottom", "top"}
    assert len(b) == int((nb < 0).sum())


def test_gather_equals_mirror_on_graded_mesh():
    mesh, alg, q, bc = S.hills_graded(length=1000.0, fine_area=40.0, coarse_area=300.0)
    runs = []
    for gather in (True, False):
        o = S.setup(OracleDomain(mesh, OracleKernels(gather=gather)), alg, q, bc)
        for _ in o.evolve(yieldstep=2.0, finaltime=6.0):
            pass
        runs.append((list(o.dt_history), o.stage_c.copy(), o.xmom_c.copy(), o.ymom_c.copy()))
    assert runs[0][0] == runs[1][0]
    for a, b in zip(runs[0][1:], runs[1][1:]):
        assert np.array_equal(a, b)

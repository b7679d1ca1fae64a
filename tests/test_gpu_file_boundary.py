"""File_boundary and Field_boundary on the device (MI355X): the host
interpolates the file's values at each substep time and uploads them per
edge; the flux kernel applies them.

* Against the oracle (the same boundary values through its own
  update_boundary): identical timestep sequences and states bit for bit, for
  a time- and space-varying field on one side, euler and rk2.
* The reference's spatio-temporal boundary check
  (test_shallow_water_domain.py:6332-6441): a run stored to SWW drives a
  second domain, cut from the first along its diagonal, through a
  Field_boundary; the second domain reaches the first's state at the shared
  triangles (rtol 1e-4).
"""

import os

import numpy as np
import pytest

from anuga_core_amd.boundaries import Dirichlet_boundary, Reflective_boundary
from anuga_core_amd.domain import Domain
from anuga_core_amd.file_boundary import Field_boundary, File_boundary
from anuga_core_amd.mesh import Mesh, rectangular, rectangular_cross
from oracle.kernels import OracleKernels
from oracle.oracle_domain import OracleDomain
from tests.test_file_boundary import _write_sww

pytestmark = pytest.mark.gpu


def _setup(dom, alg, path, kind):
    mesh = dom.mesh
    x, y = mesh.centroid_coordinates.T
    z = -0.02 * x
    q = dict(elevation=z, stage=np.maximum(0.4 + 0.0 * x, z), friction=np.full(len(x), 0.02))
    dom.set_flow_algorithm(alg)
    for name, v in q.items():
        if hasattr(dom, "set_centroid_values"):
            dom.set_centroid_values(name, v)
        else:
            dom.set_quantity(name, v, location="centroids")
    B = (File_boundary(path, dom) if kind == "file"
         else Field_boundary(path, dom, mean_stage=0.05))
    dom.set_boundary({"left": B, "right": Reflective_boundary(dom),
                      "top": Reflective_boundary(dom), "bottom": Reflective_boundary(dom)})
    return dom


@pytest.mark.parametrize("alg", ["DE0", "DE1"])
@pytest.mark.parametrize("kind", ["file", "field"])
def test_file_boundary_bitwise_vs_oracle(tmp_path, alg, kind):
    print("case file boundary %s %s" % (alg, kind), flush=True)
    src = Mesh(*rectangular_cross(6, 6, 60.0, 30.0))
    path = str(tmp_path / "src.sww")
    _write_sww(path, src.nodes, src.triangles, np.array([0.0, 3.0, 8.0, 20.0]))
    mesh = Mesh(*rectangular_cross(16, 8, 60.0, 30.0))
    d = _setup(Domain(mesh=mesh), alg, path, kind)
    o = _setup(OracleDomain(mesh, OracleKernels(gather=True)), alg, path, kind)
    for _ in d.evolve(yieldstep=2.0, finaltime=6.0):
        pass
    for _ in o.evolve(yieldstep=2.0, finaltime=6.0):
        pass
    assert len(d.dt_history) > 10
    assert d.dt_history == o.dt_history
    for name, oname in (("stage", "stage_c"), ("xmomentum", "xmom_c"), ("ymomentum", "ymom_c")):
        assert np.array_equal(d.get_quantity(name).centroid_values, getattr(o, oname)), name


def test_spatio_temporal_field_boundary(tmp_path):
    print("case spatio-temporal field boundary", flush=True)
    points, vertices, boundary = rectangular(3, 3)
    d1 = Domain(points, vertices, boundary)
    d1.set_store(True)
    d1.set_datadir(str(tmp_path))
    d1.set_name("spatio_temporal_boundary_source")
    d1.set_quantity("elevation", 0.0)
    d1.set_quantity("friction", 0.0)
    d1.set_quantity("stage", 0.0)
    Bd = Dirichlet_boundary([0.3, 0, 0])
    d1.set_boundary({"left": Bd, "top": Bd, "right": Reflective_boundary(d1),
                     "bottom": Reflective_boundary(d1)})
    for _ in d1.evolve(yieldstep=0.671, finaltime=10):
        pass
    cv1 = d1.get_quantity("stage").centroid_values.copy()
    sww = os.path.join(str(tmp_path), "spatio_temporal_boundary_source.sww")
    assert os.path.exists(sww)

    # the triangle under the diagonal of the first domain, on its nodes
    points = [[0, 0], [1.0 / 3, 0], [1.0 / 3, 1.0 / 3], [2.0 / 3, 0], [2.0 / 3, 1.0 / 3],
              [2.0 / 3, 2.0 / 3], [1, 0], [1, 1.0 / 3], [1, 2.0 / 3], [1, 1]]
    vertices = [[1, 2, 0], [3, 4, 1], [2, 1, 4], [4, 5, 2], [6, 7, 3], [4, 3, 7], [7, 8, 4],
                [5, 4, 8], [8, 9, 5]]
    boundary = {(0, 1): "bottom", (1, 1): "bottom", (4, 1): "bottom", (4, 2): "right",
                (6, 2): "right", (8, 2): "right", (0, 0): "diagonal", (3, 0): "diagonal",
                (8, 0): "diagonal"}
    d2 = Domain(points, vertices, boundary)
    d2.set_quantity("elevation", 0.0)
    d2.set_quantity("friction", 0.0)
    d2.set_quantity("stage", 0.0)
    Bf = Field_boundary(sww, d2)
    d2.set_boundary({"right": Reflective_boundary(d2), "bottom": Reflective_boundary(d2),
                     "diagonal": Bf})
    for _ in d2.evolve(yieldstep=0.0711, finaltime=10):
        pass
    cv2 = d2.get_quantity("stage").centroid_values
    assert np.allclose(np.take(cv1, (0, 8, 16)), np.take(cv2, (0, 3, 8)), rtol=1.0e-4)
    assert np.allclose(np.take(cv1, (0, 6, 12)), np.take(cv2, (0, 1, 4)), rtol=1.0e-4)
    assert np.allclose(np.take(cv1, (12, 14, 16)), np.take(cv2, (4, 6, 8)), rtol=1.0e-4)

"""Boundary conditions beyond reflective / transmissive / Dirichlet, on the
oracle (CPU): the reference's known answers of test_boundary_conditionsIII
(test_shallow_water_domain.py:756-836, Transmissive_stage_zero_momentum
beside Dirichlet and reflective edges), the time at which time-dependent
boundaries are evaluated (test_boundary_condition_time, :838-881: at a
yield the boundary holds f(t) of that time; here also inside rk2 and rk3
steps), and the geometry of the normal-momentum boundary (the boundary
momentum is the edge momentum's normal component, along the normal).
"""

import math

import numpy as np

from anuga_core_amd.boundaries import (Dirichlet_boundary, Reflective_boundary, Time_boundary,
                                       Transmissive_n_momentum_zero_t_momentum_set_stage_boundary,
                                       Transmissive_stage_zero_momentum_boundary)
from anuga_core_amd.mesh import Mesh, rectangular_cross
from oracle.kernels import OracleKernels
from oracle.oracle_domain import OracleDomain


class _Dom:
    """Time source for boundary objects built outside a Domain."""

    def __init__(self, o):
        self.o = o

    def get_time(self):
        return self.o.get_time()


def test_boundary_conditions_III_known_answers():
    points = [[0.0, 0.0], [0.0, 2.0], [2.0, 0.0], [0.0, 4.0], [2.0, 2.0], [4.0, 0.0]]
    vertices = [[1, 0, 2], [1, 2, 4], [4, 2, 5], [3, 1, 4]]
    boundary = {(0, 0): "Third", (0, 2): "First", (2, 0): "Second", (2, 1): "Second",
                (3, 1): "Second", (3, 2): "Third"}
    mesh = Mesh(points, vertices, boundary)
    o = OracleDomain(mesh, OracleKernels())
    # the reference sets vertex values; its edge values are the vertex pair means
    stage_v = np.array([[1, 2, 3], [5, 5, 5], [0, 0, 9], [-6, 3, 3]], float)
    xm_v = np.array([[1, 1, 1], [2, 2, 2], [3, 3, 3], [4, 4, 4]], float)
    ym_v = np.array([[10, 10, 10], [20, 20, 20], [30, 30, 30], [40, 40, 40]], float)
    for v, e in ((stage_v, o.stage_e), (xm_v, o.xmom_e), (ym_v, o.ymom_e)):
        E = np.stack([(v[:, 1] + v[:, 2]) / 2, (v[:, 0] + v[:, 2]) / 2,
                      (v[:, 0] + v[:, 1]) / 2], axis=1)
        e[:] = E.ravel()
    D = Dirichlet_boundary([5, 2, 1])
    T = Transmissive_stage_zero_momentum_boundary(o)
    R = Reflective_boundary(o)
    o.set_boundary({"First": D, "Second": T, "Third": R})
    o.update_boundary()
    # boundary edges in sorted (vol, edge) order: (0,0) (0,2) (2,0) (2,1) (3,1) (3,2)
    assert o.stage_b[:6].tolist() == [2.5, 5.0, 4.5, 4.5, -1.5, -1.5]
    assert o.xmom_b[:6].tolist() == [1.0, 2.0, 0.0, 0.0, 0.0, -4.0]
    assert o.ymom_b[:6].tolist() == [-10.0, 1.0, 0.0, 0.0, 0.0, 40.0]


def _timed_domain(method):
    p, v, b = rectangular_cross(5, 5)
    mesh = Mesh(p, v, b)
    o = OracleDomain(mesh, OracleKernels(), {"euler": "DE0", "rk2": "DE1", "rk3": "DE2"}[method])
    o.set_centroid_values("stage", np.full(mesh.number_of_triangles, 0.5))
    calls = []

    def f(t):
        calls.append(t)
        return [0.5 + 0.1 * t, 0.0, 0.0]
    Bt = Time_boundary(domain=_Dom(o), function=f)
    R = Reflective_boundary(o)
    o.set_boundary({"left": Bt, "right": R, "top": R, "bottom": R})
    return o, Bt, calls


def test_time_boundary_evaluated_at_substep_times():
    for method, per_step in (("euler", 1), ("rk2", 2), ("rk3", 3)):
        o, Bt, calls = _timed_domain(method)
        left = np.asarray(o.mesh.tag_boundary_cells["left"])
        for t in o.evolve(yieldstep=0.05, finaltime=0.1):
            # the boundary holds f of the time of its last evaluation
            assert o.stage_b[left[0]] == 0.5 + 0.1 * calls[-1]
        dts = o.dt_history
        t0 = 0.0
        expect = []
        for dt in dts:
            if per_step == 1:
                expect += [t0]
            elif per_step == 2:
                expect += [t0, t0 + dt]
            else:
                expect += [t0, t0 + dt, t0 + dt * 0.5]
            t0 = t0 + dt
            if t0 in (0.05, 0.1):          # the yield's own update_boundary
                expect += [t0]
        # calls[0]: the check at construction; calls[1]: the initial yield
        assert calls[2:] == expect, method
        assert math.isclose(Bt.values(0.1)[0], 0.5 + 0.1 * 0.1)


def test_normal_momentum_boundary_geometry():
    p, v, b = rectangular_cross(4, 3, 4.0, 3.0)
    mesh = Mesh(p, v, b)
    o = OracleDomain(mesh, OracleKernels())
    rng = np.random.default_rng(0)
    o.xmom_e[:] = rng.standard_normal(o.xmom_e.size)
    o.ymom_e[:] = rng.standard_normal(o.ymom_e.size)
    B = Transmissive_n_momentum_zero_t_momentum_set_stage_boundary(_Dom(o), lambda t: [1.25])
    o.set_boundary({t: B for t in ("left", "right", "top", "bottom")})
    o.update_boundary()
    nrm = o.normals.reshape(-1, 6)
    vol, edge = mesh.boundary_cells, mesh.boundary_edges
    n1, n2 = nrm[vol, 2 * edge], nrm[vol, 2 * edge + 1]
    qx = o.xmom_e.reshape(-1, 3)[vol, edge]
    qy = o.ymom_e.reshape(-1, 3)[vol, edge]
    M = mesh.boundary_length
    assert np.all(o.stage_b[:M] == 1.25)
    assert np.allclose(o.xmom_b[:M] * n1 + o.ymom_b[:M] * n2, qx * n1 + qy * n2, atol=1e-14)
    assert np.allclose(-o.xmom_b[:M] * n2 + o.ymom_b[:M] * n1, 0.0, atol=1e-14)

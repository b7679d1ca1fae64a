"""Golden trajectories held by the reference's own tests
(anuga/shallow_water/tests/test_shallow_water_domain.py), as data: each case
builds the same scenario on an oracle domain or a device Domain, evolves it
through several yields, and compares with the reference's expected arrays.

These are the fixtures that pin the multi-yield branch of _evolve_base
(generic_domain.py:1855-1911): every yield distributes, applies boundaries and
resets the recorded timestep extremes before the next step.

Expected values are the reference's (8 significant digits); the reference
checks them with numpy.allclose (rtol 1e-5, atol 1e-8), as here.
"""

from __future__ import annotations

import numpy as np

from anuga_core_amd.boundaries import Dirichlet_boundary, Reflective_boundary
from anuga_core_amd.mesh import Mesh, rectangular

BETAS = ("beta_w", "beta_w_dry", "beta_uh", "beta_uh_dry", "beta_vh", "beta_vh_dry")

# test_bedslope_problem_second_order_two_steps (:5717-5813) and
# ..._two_yieldsteps (:5815-5911): stage after evolve(0.05, 0.1)
W_TWO_STEPS = [
    0.01308246, 0.02201217, 0.01358687, 0.023575, 0.01370201, 0.0235574, 0.01370201, 0.02355753,
    0.0136998, 0.02361447, 0.01454146, 0.02507853, -0.04254524, -0.022535, -0.04260359,
    -0.0225144, -0.04263851, -0.02252512, -0.04263809, -0.02252512, -0.04264303, -0.02249883,
    -0.04257228, -0.02296247, -0.0981472, -0.07964098, -0.09814502, -0.07968513, -0.09814337,
    -0.07968807, -0.09814296, -0.07968807, -0.09814296, -0.07968807, -0.09814555, -0.07965964,
    -0.15368138, -0.13518778, -0.15366285, -0.13519037, -0.15366285, -0.13519037, -0.15366285,
    -0.13518996, -0.15366479, -0.13518832, -0.15369898, -0.13518613, -0.21012122, -0.19071462,
    -0.21066837, -0.19057453, -0.21064807, -0.19057613, -0.21064807, -0.19057519, -0.21067682,
    -0.19061141, -0.21080033, -0.19066023, -0.25755282, -0.24897294, -0.25604086, -0.24820174,
    -0.25598276, -0.24820504, -0.25598272, -0.24820504, -0.2559903, -0.2480719, -0.25395141,
    -0.24799187]

# test_bedslope_problem_second_order_more_steps (:5913-6062): evolve(0.05, 0.5)
W_MORE_STEPS = [
    -0.0301883, -0.01127593, -0.02834861, -0.0108968, -0.02806583, -0.01074475, -0.02788852,
    -0.01065176, -0.02753658, -0.01013457, -0.02602057, -0.00697983, -0.07333722, -0.06511481,
    -0.08002821, -0.06234795, -0.07884103, -0.06214974, -0.07867692, -0.06199716, -0.07848496,
    -0.06160417, -0.07782372, -0.06214378, -0.12379366, -0.10782217, -0.12402788, -0.10970479,
    -0.12407233, -0.10911802, -0.12375966, -0.10870709, -0.1233677, -0.10897925, -0.1246598,
    -0.10913556, -0.16030637, -0.15059851, -0.1619466, -0.15150383, -0.16222821, -0.15126241,
    -0.16204491, -0.15117909, -0.16170723, -0.15047256, -0.16159564, -0.15306895, -0.20650721,
    -0.19412466, -0.20661286, -0.193611, -0.20611843, -0.1932463, -0.20476096, -0.19336801,
    -0.20673157, -0.1928249, -0.20666551, -0.18982871, -0.14177628, -0.13867427, -0.138984,
    -0.13970989, -0.13837738, -0.13785944, -0.13772734, -0.13619726, -0.13708211, -0.13865006,
    -0.13413676, -0.14008116]
UH_MORE_STEPS = [
    0.00439415, 0.00080998, 0.00359639, 0.00086799, 0.00369947, 0.00089459, 0.00376748,
    0.00091612, 0.00394185, 0.00100331, 0.00423642, 0.00102681, 0.01866626, 0.00841126,
    0.01265179, 0.00989651, 0.01358962, 0.01007976, 0.01375394, 0.01018847, 0.01388854,
    0.01057534, 0.01516146, 0.01155302, 0.03130996, 0.02572364, 0.03012732, 0.02348833,
    0.02999708, 0.02419562, 0.03046757, 0.02471555, 0.03088795, 0.02415592, 0.02886438,
    0.02425477, 0.06609542, 0.04946308, 0.06236933, 0.04715456, 0.06176596, 0.04743353,
    0.06204203, 0.04760272, 0.06262964, 0.04858867, 0.06290163, 0.0447266, 0.08245285,
    0.07715851, 0.08533206, 0.07817769, 0.08610624, 0.07901377, 0.0879718, 0.07879576,
    0.08550203, 0.07963651, 0.08581205, 0.08376399, 0.01997027, 0.07948531, 0.02703306,
    0.08443647, 0.02694374, 0.08404407, 0.02627307, 0.08400949, 0.02649757, 0.08397135,
    0.03080026, 0.09688469]
VH_MORE_STEPS = [
    -4.55438497e-04, 5.60449890e-05, -1.05041759e-04, 1.09474097e-04, 3.37384149e-05,
    1.48217622e-04, 8.15006870e-05, 1.41778277e-04, -8.95699377e-05, -1.50989506e-05,
    -1.13309713e-03, -6.56027069e-04, 4.08407333e-04, -5.37824083e-04, -6.07324848e-04,
    -1.79169050e-04, -2.26068995e-04, -9.40081480e-05, -1.92103118e-04, -7.38055580e-05,
    -2.61634534e-04, -2.18534710e-04, -3.97532842e-04, -4.38793389e-04, 4.02437237e-04,
    1.71996550e-04, -4.71404571e-04, -7.07195798e-04, -2.92606384e-04, -2.89749829e-04,
    -2.14551908e-04, -3.47586675e-04, -4.51408518e-04, -2.85326203e-04, 2.60585517e-04,
    1.85996351e-04, 1.76070321e-03, 4.27131620e-04, -1.36487970e-04, -7.48486546e-04,
    -2.25542208e-04, -4.38919415e-04, -1.23010639e-05, -3.49340425e-04, -2.36098938e-04,
    -7.63395297e-04, 6.66099385e-05, 6.22983544e-04, 4.52833391e-04, 2.24808947e-03,
    4.85972740e-04, 7.50193061e-04, 4.53207970e-04, 1.05078134e-03, 6.32114704e-04,
    1.24869321e-03, 4.06577106e-04, 1.01553335e-03, -5.26250901e-04, 1.78062345e-04,
    -2.04979836e-03, -2.70644308e-03, -3.34067897e-03, -1.94716541e-03, -2.28075505e-03,
    -1.61494727e-03, -2.15457373e-03, -1.48932625e-03, -3.24804437e-03, -1.09719715e-03,
    3.38706650e-03, -8.98151209e-04]

# test_bedslope_problem_first_order_moresteps_low_froude_0 (:5505-5564) and
# ..._low_froude_1 (:5566-5624): DE0 defaults, evolve(0.05, 0.5)
W_FIRST_ORDER_MORESTEPS = [
    -0.02933401, -0.01091747, -0.02782667, -0.01057726, -0.02746728, -0.01033124, -0.02724838,
    -0.01019733, -0.02684617, -0.00965007, -0.02495054, -0.00706089, -0.06890704, -0.06303447,
    -0.07568528, -0.06161392, -0.07566147, -0.06125286, -0.07537564, -0.06094254, -0.07488436,
    -0.06058909, -0.07465553, -0.06045194, -0.12205956, -0.10701804, -0.12318423, -0.10795678,
    -0.12302672, -0.10763744, -0.12279702, -0.10731376, -0.12214772, -0.1071946, -0.12283814,
    -0.10858945, -0.16389971, -0.15304369, -0.16428248, -0.15298725, -0.16430596, -0.15254043,
    -0.16407102, -0.15236021, -0.16390702, -0.15187689, -0.16460111, -0.15549742, -0.2083898,
    -0.19912702, -0.20845637, -0.19709535, -0.20648985, -0.19663999, -0.20525695, -0.19631453,
    -0.20446298, -0.19578452, -0.20735693, -0.19565282, -0.14022868, -0.14262659, -0.13774131,
    -0.14132395, -0.13707526, -0.14041639, -0.13594765, -0.13910709, -0.13533594, -0.1393996,
    -0.1328638, -0.1363085]

# test_flatbed_first_order (:5192-5229): xmomentum edge values [:4, 0..2]
UH_EDGES_FLATBED_FIRST = [[0.07335652, 0.06685681, 0.07071273, 0.06628975],
                          [0.07343497, 0.06685681, 0.07083783, 0.06628975],
                          [0.08124162, 0.06685681, 0.07891946, 0.06628975]]

# test_flatbed_second_order (:5232-5285) / _vmax_0 (:5287-5329): vertex 0 of
# the first 4 triangles after evolve(0.01, 0.03)
FLATBED_SECOND_DT = 0.018940360210353942
W_VERTEX_FLATBED_SECOND = [-0.00524972, 0.05350326, 0.00077479, 0.05356756]
UH_VERTEX_FLATBED_SECOND = [-0.00271431, 0.02744767, 0.00023944, 0.02746294]
VH_VERTEX_FLATBED_SECOND = [1.10413075e-03, 2.62134850e-04, -2.72890315e-05, 2.77104392e-04]


class Access:
    """The same reads and writes on an oracle domain or a device Domain."""

    def __init__(self, d):
        self.d = d
        self.oracle = hasattr(d, "set_vertex_values")

    def set_vertex(self, name, vtx):
        if self.oracle:
            self.d.set_vertex_values(name, vtx)
        else:
            self.d.set_quantity(name, vtx, location="vertices")

    def centroids(self, name):
        if self.oracle:
            return self.d.get_centroid_values(name)
        return self.d.get_quantity(name).centroid_values

    def edges(self, name):
        if self.oracle:
            key = {"stage": "stage", "xmomentum": "xmom", "ymomentum": "ymom"}[name]
            return getattr(self.d, key + "_e").reshape(-1, 3).copy()
        return self.d.get_quantity(name).edge_values

    def vertices(self, name):
        if self.oracle:
            key = {"stage": "stage", "xmomentum": "xmom", "ymomentum": "ymom"}[name]
            return getattr(self.d, key + "_v").reshape(-1, 3).copy()
        return self.d.get_quantity(name).vertex_values


def _bedslope(make, betas):
    """rectangular(6,6); bed -x/3 and stage bed + 0.05 given at vertices;
    reflective all round; DE0 (the Domain default)."""
    p, v, b = rectangular(6, 6)
    mesh = Mesh(p, v, b)
    d = make(mesh)
    d.set_flow_algorithm("DE0")
    if betas is not None:
        for k in BETAS:
            setattr(d, k, betas)
    if hasattr(d, "set_low_froude"):
        d.set_low_froude(0)
    a = Access(d)
    zv = (-mesh.vertex_coordinates[:, 0] / 3).reshape(-1, 3)
    a.set_vertex("elevation", zv)
    a.set_vertex("stage", zv + 0.05)
    d.set_boundary({t: Reflective_boundary(d) for t in ("left", "right", "top", "bottom")})
    return d, a


def _flatbed(make, betas, inflow):
    """rectangular(8,8), flat dry bed, Dirichlet [inflow, 0, 0] on the left,
    reflective elsewhere, DE0."""
    p, v, b = rectangular(8, 8)
    mesh = Mesh(p, v, b)
    d = make(mesh)
    d.set_flow_algorithm("DE0")
    if betas is not None:
        for k in BETAS:
            setattr(d, k, betas)
    Br = Reflective_boundary(d)
    d.set_boundary({"left": Dirichlet_boundary([inflow, 0.0, 0.0]), "right": Br, "top": Br,
                    "bottom": Br})
    return d, Access(d)


def _run(d, yieldstep, finaltime):
    times = [t for t in d.evolve(yieldstep=yieldstep, finaltime=finaltime)]
    return times


def case_bedslope_two_steps(make):
    d, a = _bedslope(make, 0.9)
    times = _run(d, 0.05, 0.1)
    assert np.allclose(times, [0.0, 0.05, 0.1])
    assert np.allclose(d.recorded_min_timestep, 0.0344607459654)
    assert np.allclose(d.recorded_max_timestep, 0.0391090502542)
    assert np.allclose(a.centroids("stage"), W_TWO_STEPS)
    return d


def case_bedslope_more_steps(make):
    d, a = _bedslope(make, 0.9)
    times = _run(d, 0.05, 0.5)
    assert len(times) == 11
    assert np.allclose(a.centroids("stage"), W_MORE_STEPS)
    assert np.allclose(a.centroids("xmomentum"), UH_MORE_STEPS)
    assert np.allclose(a.centroids("ymomentum"), VH_MORE_STEPS)
    return d


def case_bedslope_first_order_moresteps(make):
    d, a = _bedslope(make, None)
    times = _run(d, 0.05, 0.5)
    assert len(times) == 11
    assert np.allclose(a.centroids("stage"), W_FIRST_ORDER_MORESTEPS)
    return d


def case_flatbed_first_order(make):
    d, a = _flatbed(make, None, 0.2)
    times = _run(d, 0.02, 0.5)
    assert len(times) == 26
    e = a.edges("xmomentum")
    for i in range(3):
        assert np.allclose(e[:4, i], UH_EDGES_FLATBED_FIRST[i]), i
    return d


def case_flatbed_second_order(make):
    d, a = _flatbed(make, 0.9, 0.2)
    _run(d, 0.01, 0.03)
    assert np.allclose(d.recorded_min_timestep, FLATBED_SECOND_DT)
    assert np.allclose(d.recorded_max_timestep, FLATBED_SECOND_DT)
    assert np.allclose(a.vertices("stage")[:4, 0], W_VERTEX_FLATBED_SECOND)
    assert np.allclose(a.vertices("xmomentum")[:4, 0], UH_VERTEX_FLATBED_SECOND)
    assert np.allclose(a.vertices("ymomentum")[:4, 0], VH_VERTEX_FLATBED_SECOND)
    return d


CASES = [case_bedslope_two_steps, case_bedslope_more_steps,
         case_bedslope_first_order_moresteps, case_flatbed_first_order,
         case_flatbed_second_order]

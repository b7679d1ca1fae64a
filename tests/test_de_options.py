"""DE options beyond the presets on the CPU: low_froude 1 and 2 (the local
Froude scaling of the momentum diffusion, sw_domain_orig.c:409-425) and
sloped Manning friction (sw_domain_orig.c:2840-2888 with _gradient,
anuga/utilities/util_ext.h:52-92). The oracle's edge-once loop equals the
reference's compiled C bit for bit over whole runs; sloped friction also in
the per-triangle loop; and closed forms pin what the compiled C alone
leaves open (a flat bed makes sloped friction the flat one; on a plane the
slope factor is sqrt(1 + |grad z|^2)).
"""

import math
import os

import numpy as np
import pytest

from anuga_core_amd.mesh import Mesh, rectangular_cross
from oracle.kernels import HERE as ORACLE_DIR
from oracle.kernels import OracleKernels, RefKernels
from oracle.oracle_domain import OracleDomain
from tests import scenarios as S

HAVE_REF = os.path.exists(os.path.join(ORACLE_DIR, "_ref", "libref_sw.so"))
need_ref = pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built (no reference tree)")


def options_run(K, builder, kw, ys, ft, low_froude=0, sloped=False):
    mesh, alg, q, bc = builder(**kw)
    d = S.setup(OracleDomain(mesh, K), alg, q, bc)
    d.low_froude = low_froude
    d.use_sloped_mannings = sloped
    for _ in d.evolve(yieldstep=ys, finaltime=ft):
        pass
    return d


CASES = [(S.hills_de1, dict(m=12, n=12, seed=3), 5.0, 25.0),
         (S.dam_break_de0, dict(m=16, n=12), 1.0, 4.0)]


@need_ref
@pytest.mark.parametrize("builder,kw,ys,ft", CASES)
@pytest.mark.parametrize("low_froude,sloped", [(1, False), (2, False), (0, True), (2, True)])
def test_options_bitwise_vs_reference_c(builder, kw, ys, ft, low_froude, sloped):
    r = options_run(RefKernels(), builder, kw, ys, ft, low_froude, sloped)
    loops = [False] if low_froude else [False, True]
    for gather in loops:
        o = options_run(OracleKernels(gather=gather), builder, kw, ys, ft,
                        low_froude, sloped)
        assert r.dt_history == o.dt_history
        for a in ("stage_c", "xmom_c", "ymom_c"):
            assert np.array_equal(getattr(r, a), getattr(o, a)), (a, gather)


def test_options_change_the_run():
    """Each option acts: the runs differ from the plain preset."""
    base = options_run(OracleKernels(), S.hills_de1, dict(m=10, n=10, seed=3), 5.0, 15.0)
    for lf, sl in ((1, False), (2, False), (0, True)):
        o = options_run(OracleKernels(), S.hills_de1, dict(m=10, n=10, seed=3), 5.0, 15.0, lf, sl)
        assert not np.array_equal(o.xmom_c, base.xmom_c), (lf, sl)


def _friction_pair(z_fn, n=0.03):
    p, v, b = rectangular_cross(4, 4, 40.0, 40.0)
    mesh = Mesh(p, v, b)
    N = mesh.number_of_triangles
    k = OracleKernels()
    d = OracleDomain(mesh, k, "DE1")
    V = mesh.vertex_coordinates.reshape(N, 3, 2)
    zv = z_fn(V[:, :, 0], V[:, :, 1])
    d.elev_v[:] = zv.ravel()
    d.elev_c[:] = zv.mean(axis=1)
    d.stage_c[:] = d.elev_c + 2.0
    d.xmom_c[:] = 0.5
    d.ymom_c[:] = -0.25
    d.friction_c[:] = n
    out = []
    for sloped in (False, True):
        d.use_sloped_mannings = sloped
        d.xmom_si[:] = 0.0
        d.ymom_si[:] = 0.0
        k.friction()
        out.append(d.xmom_si.copy())
    return out, zv


def test_sloped_friction_closed_forms():
    # flat bed: the slope factor is 1 and the mean vertex bed the centroid bed
    (flat, sloped), _ = _friction_pair(lambda x, y: 0.0 * x - 3.0)
    assert np.array_equal(flat, sloped)
    # a plane z = 0.2 x - 0.1 y: factor sqrt(1 + 0.04 + 0.01) everywhere
    (flat, sloped), _ = _friction_pair(lambda x, y: 0.2 * x - 0.1 * y)
    assert np.allclose(sloped / flat, math.sqrt(1.05), rtol=1e-13)
    # the closed form itself: -g n^2 zs |q| / h^(7/3) * uh
    g, n, h, uh, vh = 9.8, 0.03, 2.0, 0.5, -0.25
    expect = -g * n * n * math.sqrt(1.05) * math.sqrt(uh * uh + vh * vh) / h ** (7.0 / 3.0) * uh
    assert np.allclose(sloped, expect, rtol=1e-13)

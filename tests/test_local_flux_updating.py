"""Local extrapolation and flux updating (max_flux_update_frequency > 1) on
the CPU: the oracle's restatement against the reference's compiled C, the
reference's own acceptance test, and the order-free form of the frequency
update that the device kernels use.

* oracle == reference C (oracle/_ref, a private copy per run so that its
  static counters start fresh): identical dt sequences and states bit for
  bit, DE0 with 2**3 and 2**8 levels, on a structured and an unstructured
  mesh (sw_domain_orig.c:509-669, :765-1150).
* test_local_extrapolation_and_flux_updating.py:61-88 on the oracle: the
  volume within 5% of the every-step run, the boundary flux integral equal
  to the volume for both, stages within 0.02 (the reference meshes the square
  with `triangle`, absent here; rectangular_cross and graded_square stand in).
* The frequency update restated per half-edge (each half of an edge replays
  the reference's two sequential visits of the pair on its own copy) gives
  the sequential loop's frequencies, flags and already_computed_flux after
  every step of a run.
* rk2 is refused with the reference's exception.
"""

import os

import numpy as np
import pytest

from oracle.kernels import OracleKernels, RefKernels
from oracle.oracle_domain import OracleDomain
from tests import scenarios as S

HAVE_REF = os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref",
                                       "libref_sw.so"))


def _run(kernels, graded, nlevels, finaltime, yieldstep=0.5, fine=False):
    mesh, alg, q, bc = (S.inflow_over_slope(24, 24, graded=graded, fine_area=8.0) if fine
                        else S.inflow_over_slope(graded=graded))
    d = S.setup(OracleDomain(mesh, kernels), alg, q, bc)
    if nlevels is not None:
        d.set_local_extrapolation_and_flux_updating(nlevels=nlevels)
    for _ in d.evolve(yieldstep=yieldstep, finaltime=finaltime):
        pass
    return d


@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref is not built (needs the reference tree)")
@pytest.mark.parametrize("graded", [False, True])
@pytest.mark.parametrize("nlevels", [3, 8])
def test_oracle_matches_reference_c(graded, nlevels):
    o = _run(OracleKernels(), graded, nlevels, 12.0, yieldstep=2.0, fine=True)
    r = _run(RefKernels(fresh=True), graded, nlevels, 12.0, yieldstep=2.0, fine=True)
    assert len(o.dt_history) > 50
    assert o.dt_history == r.dt_history
    for a in ("stage_c", "xmom_c", "ymom_c"):
        assert np.array_equal(getattr(o, a), getattr(r, a)), a
    assert o.boundary_flux_integral == r.boundary_flux_integral
    # fluxes were skipped: fewer distinct timesteps than steps would show a
    # frozen timestep; the timestep may grow only every 2**nlevels steps
    dts = np.array(o.dt_history)
    grows = np.nonzero(dts[1:] > dts[:-1] * (1 + 1e-12))[0] + 1
    assert len(grows) < len(dts) // 2


@pytest.mark.parametrize("graded", [False, True])
def test_reference_acceptance_de0(graded):
    d1 = _run(OracleKernels(), graded, None, 20.0, yieldstep=0.1)
    d2 = _run(OracleKernels(), graded, 8, 20.0, yieldstep=0.1)
    v1, v2 = d1.get_water_volume(), d2.get_water_volume()
    assert v1 > 0.0
    assert np.allclose(v1, v2, rtol=0.05)
    assert np.allclose(v1, d1.boundary_flux_integral)
    assert np.allclose(v2, d2.boundary_flux_integral)
    assert np.all(np.abs(d1.stage_c - d2.stage_c) < 0.02)
    # the timestep is the yieldstep throughout (as in the reference's test),
    # so only the skipped fluxes separate the runs
    assert not np.array_equal(d1.stage_c, d2.stage_c)


def test_rk2_is_refused():
    mesh, alg, q, bc = S.inflow_over_slope()
    d = S.setup(OracleDomain(mesh, OracleKernels()), "DE1", q, bc)
    with pytest.raises(Exception, match="euler"):
        d.set_local_extrapolation_and_flux_updating(nlevels=8)


def _frequency_per_half_edge(nb, nbe, et, fuf, acf, cyc, M, timestep):
    """The frequency update with each half-edge on its own (the device's
    form): returns new (fuf, unf, acf)."""
    N = len(nb)
    fuf, acf = fuf.copy(), acf.copy()

    def derive(a, b, pair):
        f = int(min((a / timestep) * 1.0, M * 1.0))
        if pair:
            f = int(min(float(int(min(b / timestep * 1.0, M * 1.0))), float(f)))
        f = min(max(f, 1), M)
        ii = 2
        while ii < M:
            if ii < f < 2 * ii:
                f = ii
            ii *= 2
        return f

    tmin = np.empty(N, np.int64)
    for k in range(N):
        for i in range(3):
            ki = 3 * k + i
            f = int(fuf[ki])
            if cyc & (f - 1) == 0:
                a = int(acf[ki])
                n = nb[k, i]
                if n >= 0:
                    go = a != 0
                else:
                    go = a > 0
                    a = -a
                if go:
                    nm = 3 * n + nbe[k, i]
                    f = derive(et[ki], et[nm] if n >= 0 else 0.0, n >= 0)
                    if n >= 0 and a > 0 and cyc & (f - 1) != 0:
                        a = -a
                    fuf[ki] = f
                acf[ki] = a
        tmin[k] = fuf[3 * k:3 * k + 3].min()
    unf = np.empty(3 * N, np.int64)
    for k in range(N):
        for i in range(3):
            n = nb[k, i]
            f = min(tmin[k], tmin[n]) if n >= 0 else tmin[k]
            fuf[3 * k + i] = f
            unf[3 * k + i] = 1 if (cyc + 1) & (f - 1) == 0 else 0
    return fuf, unf, acf


@pytest.mark.parametrize("graded", [False, True])
def test_per_half_edge_frequency_update_matches_the_loop(graded):
    mesh, alg, q, bc = S.inflow_over_slope(24, 24, graded=graded, fine_area=8.0)
    k = OracleKernels()
    d = S.setup(OracleDomain(mesh, k), alg, q, bc)
    M = 2 ** 3
    d.set_local_extrapolation_and_flux_updating(nlevels=3)
    nb, nbe = mesh.neighbours, mesh.neighbour_edges
    a = k._lts_arrays
    checked = []
    seen_neg = False
    orig = k.compute_flux_update_frequency

    def checked_update(timestep):
        before = (a["fuf"].copy(), a["acf"].copy(), d.edge_timestep.copy())
        cyc = (k.lts.cyclic_number_of_steps + 1) % M
        orig(timestep)
        fuf, unf, acf = _frequency_per_half_edge(nb, nbe, before[2], before[0], before[1],
                                                 cyc, M, timestep)
        assert np.array_equal(fuf, a["fuf"])
        assert np.array_equal(unf, a["unf"])
        assert np.array_equal(acf, a["acf"])
        checked.append(int((a["fuf"] > 1).sum()))

    k.compute_flux_update_frequency = checked_update
    for _ in d.evolve(yieldstep=2.0, finaltime=6.0):
        seen_neg = seen_neg or bool((a["acf"] < 0).any())
    assert len(checked) > 30
    assert max(checked) > 0 and seen_neg   # frequencies above 1 and flipped signs occurred

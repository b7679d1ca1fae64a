"""Ensembles on the device (MI355X): every scenario of an ensemble handle must
evolve bit for bit as a single-scenario Domain of the same inputs does (the
Domain itself is pinned to the oracle by the parity tests), with its own
timestep sequence, while the scenarios share every launch."""

import numpy as np
import pytest

from anuga_core_amd.domain import Domain
from anuga_core_amd.ensemble import Ensemble
from anuga_core_amd.mesh import Mesh, rectangular, rectangular_cross
from tests import scenarios as S

pytestmark = pytest.mark.gpu


def _scenarios(seed=3):
    """Scenarios of different sizes, shapes, beds, friction and boundaries,
    sized so that several share a workgroup-padded layout."""
    rng = np.random.default_rng(seed)
    out = []
    for k, kind in ((3, "cross"), (7, "cross"), (12, "rect"), (20, "cross"), (5, "cross")):
        L = 10.0 * k
        p, v, b = (rectangular_cross if kind == "cross" else rectangular)(k, k, L, L)
        mesh = Mesh(p, v, b)
        x, y = mesh.centroid_coordinates.T
        z = -0.005 * x
        for _ in range(3):
            a, cx, cy, s = rng.uniform(0.2, 1.5), *rng.uniform(0, L, 2), rng.uniform(0.1, 0.3) * L
            z = z + a * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * s * s))
        hx, hy = rng.uniform(0.3, 0.7, 2) * L
        hump = rng.uniform(0.5, 2.0) * np.exp(-((x - hx) ** 2 + (y - hy) ** 2) / (2 * (0.15 * L) ** 2))
        w = np.maximum(hump, z)
        q = dict(elevation=z, stage=w, friction=np.full(len(x), rng.uniform(0.01, 0.06)))
        bc = {"left": "transmissive", "right": "reflective", "top": "reflective",
              "bottom": ("dirichlet", [0.1, 0.0, 0.0]) if k == 12 else "reflective"}
        out.append((mesh, q, bc))
    return out


def _ensemble(scen, alg, precision="fp64"):
    ens = Ensemble([m for m, _, _ in scen], flow_algorithm=alg, precision=precision)
    for s, (mesh, q, bc) in enumerate(scen):
        ens.set_scenario(s, mesh, q, bc)
    return ens


def _domain(mesh, q, bc, alg, precision="fp64", reorder=True):
    # the ensemble's internal order (Hilbert tiles per scenario) by default, so
    # that per-block partial sums (the boundary flux integral) add in the same
    # order
    return S.setup(Domain(mesh=mesh, precision=precision, reorder=reorder), alg, q, bc)


@pytest.mark.parametrize("alg,precision", [("DE1", "fp64"), ("DE0", "fp64"), ("DE1", "mixed")])
def test_ensemble_steps_match_single_domains(alg, precision):
    print("ensemble steps", alg, precision, flush=True)
    scen = _scenarios()
    ens = _ensemble(scen, alg, precision)
    ens.evolve_steps(150)
    for s, (mesh, q, bc) in enumerate(scen):
        d = _domain(mesh, q, bc, alg, precision)
        d.evolve_steps(150)
        assert ens.number_of_steps[s] == d.number_of_steps == 150
        assert ens.get_time(s) == d.get_time()
        for name in ("stage", "xmomentum", "ymomentum"):
            a = ens.get_quantity(s, name)
            b = d.get_quantity(name).centroid_values
            assert np.array_equal(a, b), (s, name, np.max(np.abs(a - b)))
        assert ens.get_boundary_flux_integrals()[s] == d.boundary_flux_integral
        vol = ens.get_water_volumes()[s]
        assert vol == pytest.approx(d.get_water_volume(), rel=1e-14, abs=0.0)
        # Reverse Cuthill-McKee order (the separate kernels): the same state
        # bits; the boundary flux is summed in another order
        t = _domain(mesh, q, bc, alg, precision, reorder="rcm")
        t.evolve_steps(150)
        for name in ("stage", "xmomentum", "ymomentum"):
            assert np.array_equal(t.get_quantity(name).centroid_values,
                                  d.get_quantity(name).centroid_values), name
        assert t.boundary_flux_integral == pytest.approx(d.boundary_flux_integral, rel=1e-12,
                                                         abs=1e-12)
        d.close()
        t.close()
    # the scenarios kept their own timesteps
    assert len(set(ens.get_times().tolist())) == len(scen)


def test_ensemble_evolve_yields_match_single_domains():
    print("ensemble evolve yields", flush=True)
    scen = _scenarios(seed=5)[:3]
    ens = _ensemble(scen, "DE0")
    snaps = []
    for t in ens.evolve(yieldstep=0.25, finaltime=1.0):
        snaps.append((t, [ens.get_quantity(s, "stage") for s in range(len(scen))],
                      ens.get_times().copy()))
    assert [round(t, 12) for t, _, _ in snaps] == [0.0, 0.25, 0.5, 0.75, 1.0]
    for s, (mesh, q, bc) in enumerate(scen):
        d = _domain(mesh, q, bc, "DE0")
        ys = []
        for t in d.evolve(yieldstep=0.25, finaltime=1.0):
            ys.append((t, d.get_quantity("stage").centroid_values, d.get_time()))
        assert len(ys) == len(snaps)
        for (t, w, ts), (td, wd, tdd) in zip(snaps, ys):
            assert ts[s] == tdd
            assert np.array_equal(w[s], wd)
        d.close()


def test_ensemble_volume_and_independence():
    """A closed (all reflective) scenario conserves its volume whatever its
    neighbours in the ensemble do; a scenario's result does not depend on
    which other scenarios share the handle."""
    print("ensemble volume / independence", flush=True)
    scen = _scenarios(seed=7)
    mesh, q, bc = scen[3]
    closed = (mesh, q, {t: "reflective" for t in bc})
    big = _ensemble([scen[0], closed, scen[2], scen[4]], "DE1")
    v0 = big.get_water_volumes()[1]
    big.evolve_steps(100)
    v1 = big.get_water_volumes()[1]
    assert abs(v1 - v0) <= 1e-12 * v0
    alone = _ensemble([closed], "DE1")
    alone.evolve_steps(100)
    assert np.array_equal(alone.get_quantity(0, "stage"), big.get_quantity(1, "stage"))
    assert alone.get_time(0) == big.get_time(1)

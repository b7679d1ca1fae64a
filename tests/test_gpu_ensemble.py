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
    # the ensemble's internal order (bisection tiles per scenario) by default, so
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


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_reduced_batched_workload_vs_oracle(precision):
    """The batched workload's generator at reduced size: 64 scenarios, k
    log-uniform in [8, 40] (seed 1; 256 to 6400 triangles each), DE1, one
    handle, 120 steps with per-scenario timesteps. A sample of 12 scenarios
    (the smallest, the largest and 10 drawn at random) is run on the oracle:
    FP64 gives the oracle's bits and its time (the sum of the scenario's own
    timesteps); mixed stays within the stated error (stage L1 over the water
    column <= 1e-5) and its time within 1e-4 relative."""
    import bench
    print("ensemble reduced batched workload vs oracle", precision, flush=True)
    from oracle.kernels import OracleKernels
    from oracle.oracle_domain import OracleDomain
    spec = bench.build_ensemble(64, seed=1, kmin=8, kmax=40)
    scen = [bench.ensemble_scenario(k, prm) for k, prm in spec]
    ks = np.array([k for k, _ in spec])
    assert ks.min() < 12 and ks.max() > 30
    ens = Ensemble([m for m, _, _ in scen], flow_algorithm="DE1", precision=precision)
    for s, (mesh, q, bc) in enumerate(scen):
        ens.set_scenario(s, mesh, q, bc)
    nsteps = 120
    ens.evolve_steps(nsteps)
    assert list(ens.number_of_steps) == [nsteps] * len(scen)
    rng = np.random.default_rng(11)
    sample = sorted({int(np.argmin(ks)), int(np.argmax(ks))} |
                    set(rng.choice(len(scen), 10, replace=False).tolist()))
    for s in sample:
        mesh, q, bc = scen[s]
        o = S.setup(OracleDomain(mesh, OracleKernels(gather=True)), "DE1", q, bc)
        for _ in o.evolve(yieldstep=1e300, finaltime=1e300, max_steps=nsteps):
            pass
        assert len(o.dt_history) == nsteps
        t_oracle = o.get_time()
        if precision == "fp64":
            assert ens.get_time(s) == t_oracle, (s, ens.get_time(s), t_oracle)
            for name, oname in (("stage", "stage_c"), ("xmomentum", "xmom_c"),
                                ("ymomentum", "ymom_c")):
                assert np.array_equal(ens.get_quantity(s, name), getattr(o, oname)), (s, name)
        else:
            w = ens.get_quantity(s, "stage")
            h = np.maximum(o.stage_c - o.elev_c, 0.0)
            l1 = np.sum(np.abs(w - o.stage_c) * mesh.areas) / max(np.sum(h * mesh.areas), 1e-30)
            print("scenario %d k=%d stage L1/water column %.2e" % (s, ks[s], l1), flush=True)
            assert l1 <= 1e-5, (s, l1)
            assert abs(ens.get_time(s) - t_oracle) <= 1e-4 * t_oracle
    ens.close()

"""Gradients on the device (MI355X): the discrete adjoint of a fixed-timestep
DE1 run (anuga_adjoint_record / anuga_adjoint_gradient) against

* the CPU driver of the same cell math (tests/adjoint_harness.cpp, itself
  checked against the oracle and finite differences): the recorded gauge
  series and the gradients bit for bit (the device sums every adjoint in a
  fixed order), and the same bits on a second run;
* central finite differences of the device forward and of the CPU oracle
  along random directions (steps 1e-6 .. 1e-9; see test_adjoint_cpu.py for
  why a step may straddle a kink of the limiter);
* the torch.autograd.Function built on it.
"""

import ctypes as C

import numpy as np
import pytest
import torch  # noqa: F401  (before the device library: one HIP runtime per process)

from anuga_core_amd.adjoint import GaugeRun
from anuga_core_amd.domain import Domain
from oracle.kernels import OracleKernels
from oracle.oracle_domain import OracleDomain
from tests import scenarios as S
from tests.test_adjoint_cpu import (Layout, _agrees, _wet_case, harness,  # noqa: F401
                                    smooth_direction)

pytestmark = pytest.mark.gpu


def _initial(mesh, q, seed=1):
    rng = np.random.default_rng(seed)
    x, y = mesh.centroid_coordinates.T
    N = len(x)
    return dict(stage=q["stage"].copy(),
                xmomentum=0.3 * np.sin(x / 15.0) * np.cos(y / 20.0) + 1e-4 * rng.random(N),
                ymomentum=0.2 * np.cos(x / 25.0) * np.sin(y / 10.0) + 1e-4 * rng.random(N))


def _domain(mesh, q, bc, init):
    d = S.setup(Domain(mesh=mesh), "DE1", q, bc)
    for name, v in init.items():
        d.set_quantity(name, v, location="centroids")
    return d


def _harness_run(lib, mesh, q, bc, init, dt, nsteps, gauges, obs_every, dL=None):
    L = Layout(mesh, q, bc)
    N = L.N
    S0 = np.stack([init["stage"], init["xmomentum"], init["ymomentum"], q["elevation"]], axis=1)
    S0 = np.ascontiguousarray(S0)
    g = np.ascontiguousarray(gauges, np.int32)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    Sx = S0.copy()
    rec = np.zeros((nsteps // obs_every, len(g)))
    lib.harness_forward(*L.args(), C.c_double(dt), nsteps, dp(Sx), len(g),
                        g.ctypes.data_as(C.POINTER(C.c_int)), obs_every, dp(rec))
    if dL is None:
        return rec, Sx
    nbar, zbar, lam = np.zeros(N), np.zeros(N), np.zeros(3 * N)
    lib.harness_adjoint(*L.args(), C.c_double(dt), nsteps, dp(S0), len(g),
                        g.ctypes.data_as(C.POINTER(C.c_int)), obs_every,
                        dp(np.ascontiguousarray(dL)), dp(nbar), dp(zbar), dp(lam))
    return rec, Sx, nbar, zbar, lam.reshape(3, N)


def test_gpu_adjoint_matches_cpu_driver(harness):
    print("adjoint vs CPU driver", flush=True)
    mesh, q, bc = _wet_case(9)
    init = _initial(mesh, q)
    rng = np.random.default_rng(3)
    gauges = rng.choice(mesh.number_of_triangles, 8, replace=False)
    dt, nsteps, obs_every = 0.15, 30, 5
    assert dt < 0.9 * _cfl_limit(mesh, q, bc, init)
    d = _domain(mesh, q, bc, init)
    rec = d.adjoint_record(nsteps, dt, gauges, obs_every, checkpoint_every=7)
    dL = rng.standard_normal(rec.shape)
    gn, gz, gs = d.adjoint_gradient(dL)
    hrec, hS, hn, hz, hs = _harness_run(harness, mesh, q, bc, init, dt, nsteps, gauges,
                                        obs_every, dL)
    assert np.array_equal(rec, hrec)
    assert np.array_equal(d.get_quantity("stage").centroid_values, hS[:, 0])
    for a, b in ((gn, hn), (gz, hz), (gs, hs)):
        assert np.array_equal(a, b)
    assert np.abs(gn).max() > 0 and np.abs(gz).max() > 0
    # a second run: the same bits
    d.close()
    d = _domain(mesh, q, bc, init)
    d.adjoint_record(nsteps, dt, gauges, obs_every, checkpoint_every=5)
    for a, b in zip(d.adjoint_gradient(dL), (gn, gz, gs)):
        assert np.array_equal(a, b)


def _cfl_limit(mesh, q, bc, init):
    """The oracle's CFL timestep (CFL 1, DE1) of the initial state."""
    o = OracleDomain(mesh, OracleKernels(gather=True), "DE1")
    for name in ("elevation", "friction"):
        o.set_centroid_values(name, q[name])
    for name, v in init.items():
        o.set_centroid_values(name, v)
    o.set_boundary({t: S.make_boundary(k, o) for t, k in bc.items()})
    o.distribute_to_vertices_and_edges()
    o.update_boundary()
    o.compute_fluxes(0)
    return o.flux_timestep


def test_gpu_adjoint_finite_differences_device_and_oracle():
    """A 16x16 run at dt 0.08, below its CFL limit (0.095 at the start and
    over the whole run, checked on the CPU, where the same cell math runs
    in bounds under AddressSanitizer even for a run that blows up)."""
    print("adjoint vs finite differences", flush=True)
    mesh, q, bc = _wet_case(16)
    init = _initial(mesh, q)
    N = mesh.number_of_triangles
    rng = np.random.default_rng(4)
    gauges = rng.choice(N, 12, replace=False)
    dt, nsteps, obs_every = 0.08, 60, 10
    assert dt < 0.9 * _cfl_limit(mesh, q, bc, init)
    d = _domain(mesh, q, bc, init)
    run = GaugeRun(d, init, nsteps, dt, gauges, obs_every, checkpoint_every=16)
    n0, z0 = q["friction"].copy(), q["elevation"].copy()
    rec = run.forward(n0, z0)
    obs = rec + 0.01 * rng.standard_normal(rec.shape)
    gn, gz, _ = d.adjoint_gradient(2.0 * (rec - obs))

    def loss_dev(n, z):
        return float(np.sum((run.forward(n, z) - obs) ** 2))

    def loss_oracle(n, z):
        o = OracleDomain(mesh, OracleKernels(gather=True), "DE1")
        o.set_centroid_values("elevation", z)
        o.set_centroid_values("friction", n)
        for name, v in init.items():
            o.set_centroid_values(name, v)
        o.set_boundary({t: S.make_boundary(k, o) for t, k in bc.items()})
        # the steps of _evolve_base without the initial distribute (the
        # device record starts from the raw centroid state too)
        o.fixed_flux_timestep = dt
        o._order_ = o.default_order
        o.relative_finaltime = None
        o.relative_yieldtime = 1e300
        o.relative_time = 0.0
        o.recorded_min_timestep = o.evolve_max_timestep
        o.recorded_max_timestep = o.evolve_min_timestep
        r = []
        for s in range(1, nsteps + 1):
            t0 = o.relative_time
            o.evolve_one_rk2_step(1e300, None)
            assert o.timestep == dt
            o.relative_time = t0 + o.timestep
            if s % obs_every == 0:
                r.append(o.stage_c[gauges].copy())
        return float(np.sum((np.array(r) - obs) ** 2))

    assert loss_oracle(n0, z0) == loss_dev(n0, z0)    # the same forward, bit for bit
    # friction along white-noise directions, the bed along smooth random
    # ones (smooth_direction: a white-noise bed switches limiter branches at
    # every step size down to rounding)
    for trial in range(3):
        dn = rng.standard_normal(N)
        _agrees(np.dot(gn, dn), lambda e: loss_dev(n0 + e * dn, z0))
        dz = smooth_direction(mesh, rng)
        _agrees(np.dot(gz, dz), lambda e: loss_dev(n0, z0 + e * dz))
    dn = rng.standard_normal(N)
    _agrees(np.dot(gn, dn), lambda e: loss_oracle(n0 + e * dn, z0))
    dz = smooth_direction(mesh, rng)
    _agrees(np.dot(gz, dz), lambda e: loss_oracle(n0, z0 + e * dz))


def test_torch_autograd_function():
    import torch
    print("adjoint torch autograd", flush=True)
    mesh, q, bc = _wet_case(9)
    init = _initial(mesh, q)
    gauges = np.arange(0, mesh.number_of_triangles, 37)
    d = _domain(mesh, q, bc, init)
    assert 0.15 < 0.9 * _cfl_limit(mesh, q, bc, init)
    run = GaugeRun(d, init, 20, 0.15, gauges, obs_every=4, checkpoint_every=8)
    n = torch.tensor(q["friction"], dtype=torch.float64, requires_grad=True)
    z = torch.tensor(q["elevation"], dtype=torch.float64, requires_grad=True)
    stage = run(n, z)
    obs = stage.detach() + 0.01
    loss = ((stage - obs) ** 2).sum()
    loss.backward()
    rec = run.forward(q["friction"], q["elevation"])
    gn, gz, _ = d.adjoint_gradient(2.0 * (rec - obs.numpy()))
    assert np.array_equal(n.grad.numpy(), gn) or np.allclose(n.grad.numpy(), gn, rtol=1e-12)
    assert np.allclose(z.grad.numpy(), gz, rtol=1e-12, atol=1e-15)


def test_torch_autograd_two_forwards_before_backward():
    """Each backward differentiates its own forward's run, also when another
    forward (other parameters) was recorded in between."""
    import torch
    print("adjoint torch autograd, two forwards", flush=True)
    mesh, q, bc = _wet_case(9)
    init = _initial(mesh, q)
    gauges = np.arange(0, mesh.number_of_triangles, 29)
    d = _domain(mesh, q, bc, init)
    run = GaugeRun(d, init, 12, 0.15, gauges, obs_every=3, checkpoint_every=5)
    n0 = q["friction"]
    n1 = q["friction"] * 1.5
    na = torch.tensor(n0, dtype=torch.float64, requires_grad=True)
    nb = torch.tensor(n1, dtype=torch.float64, requires_grad=True)
    z = torch.tensor(q["elevation"], dtype=torch.float64)
    sa = run(na, z)
    sb = run(nb, z)
    loss = (sa ** 2).sum() + (sb ** 2).sum()
    loss.backward()
    for n, t in ((n0, na), (n1, nb)):
        rec = run.forward(n, q["elevation"])
        gn, _, _ = d.adjoint_gradient(2.0 * rec)
        assert np.allclose(t.grad.numpy(), gn, rtol=1e-12, atol=1e-18)
    assert not np.allclose(na.grad.numpy(), nb.grad.numpy())

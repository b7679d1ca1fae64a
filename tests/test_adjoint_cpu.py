"""The generic cell math behind the discrete adjoint (anuga_core_amd/csrc/
de_adjoint.h), driven on the CPU by tests/adjoint_harness.cpp (CPU only):

* its double instantiation is the DE1 step: a fixed-timestep run equals the
  oracle's (set_fixed_flux_timestep semantics) bit for bit, wet/dry included;
* its dual-number adjoint is the derivative of that run: gradients of a gauge
  loss with respect to the friction, the bed and the initial state agree with
  finite differences along random directions.

The scheme is piecewise smooth: the gradient limiter (sw_domain_orig.c:1306-1333)
and the wet/dry and upwind choices take min / max of neighbour values, so the
run has kinks, and a central difference whose step straddles one mixes the two
one-sided slopes. The adjoint is the derivative of the piece the forward is on.
Kinks are dense in a run of many steps, so a direction passes when the
central difference at one of the steps 1e-6 .. 1e-9 agrees to 2e-5 (a smaller
step straddles fewer kinks; below 1e-9 rounding dominates).
"""

import ctypes as C
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from anuga_core_amd.mesh import Mesh, rectangular, rectangular_cross
from oracle.kernels import OracleKernels
from oracle.oracle_domain import OracleDomain
from tests import scenarios as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DE1 = np.array([9.8, 1.0e-12, 1.0e-5, 1.0, 0.0, 1.0, 0.0, 1.0, 0.0])


@pytest.fixture(scope="module")
def harness():
    tmp = tempfile.mkdtemp()
    so = os.path.join(tmp, "libharness.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-shared", "-fPIC",
                    "-I" + os.path.join(ROOT, "anuga_core_amd", "csrc"), "-o", so,
                    os.path.join(ROOT, "tests", "adjoint_harness.cpp")], check=True)
    return C.CDLL(so)


class Layout:
    """The library's per-triangle layouts, in the mesh's own order."""

    def __init__(self, mesh, q, bc):
        N = self.N = mesh.number_of_triangles
        self.M = mesh.boundary_length
        nbr, ne = mesh.neighbours, mesh.neighbour_edges
        nb = np.zeros((N, 4), np.int32)
        for i in range(3):
            nb[:, i] = np.where(nbr[:, i] >= 0, (nbr[:, i] << 2) | ne[:, i], nbr[:, i])
        nb[:, 3] = mesh.number_of_boundaries
        self.nb = nb
        cen = np.ascontiguousarray(mesh.centroid_coordinates)
        em = mesh.edge_midpoint_coordinates.reshape(N, 3, 2)
        self.cen = cen
        self.gx = np.ascontiguousarray(np.concatenate([em[:, :, 0] - cen[:, 0:1],
                                                       em[:, :, 1] - cen[:, 1:2]], axis=1))
        nrm = mesh.normals.reshape(N, 3, 2)
        self.nx = np.ascontiguousarray(nrm[:, :, 0].T)
        self.ny = np.ascontiguousarray(nrm[:, :, 1].T)
        self.len = np.ascontiguousarray(mesh.edgelengths.reshape(N, 3).T)
        self.area = np.ascontiguousarray(mesh.areas)
        self.eta = np.ascontiguousarray(q["friction"], np.float64).copy()
        # boundary kinds of the edge function: 0 reflective, 1 transmissive,
        # 2 Dirichlet ([stage, xmom, ymom] in bval), 3 centroid-transmissive
        kinds = {"reflective": 0, "transmissive": 1, "transmissive_centroid": 3}
        Mb = max(self.M, 1)
        self.bkind = np.zeros(Mb, np.int8)
        self.bval = np.zeros(3 * Mb)
        for tag, ids in mesh.tag_boundary_cells.items():
            ids = np.asarray(ids)
            if isinstance(bc[tag], tuple):
                self.bkind[ids] = 2
                for c in range(3):
                    self.bval[c * Mb + ids] = bc[tag][1][c]
            else:
                self.bkind[ids] = kinds[bc[tag]]
        self.S = np.zeros((N, 4))
        self.S[:, 0] = q["stage"]
        self.S[:, 3] = q["elevation"]

    def args(self):
        p = lambda a, t: a.ctypes.data_as(C.POINTER(t))
        d = C.c_double
        return [self.N, self.M, p(self.nb, C.c_int), p(self.cen, d), p(self.gx, d), p(self.nx, d),
                p(self.ny, d), p(self.len, d), p(self.area, d), p(self.eta, d),
                p(self.bkind, C.c_byte), p(self.bval, d), p(DE1, d)]


def _forward(lib, L, S0, dt, nsteps, gauges, obs_every):
    S = np.ascontiguousarray(S0, np.float64).copy()
    g = np.ascontiguousarray(gauges, np.int32)
    rec = np.zeros((max(nsteps // obs_every, 1), len(g)))
    lib.harness_forward(*L.args(), C.c_double(dt), nsteps,
                        S.ctypes.data_as(C.POINTER(C.c_double)), len(g),
                        g.ctypes.data_as(C.POINTER(C.c_int)), obs_every,
                        rec.ctypes.data_as(C.POINTER(C.c_double)))
    return S, rec


def test_forward_equals_oracle_fixed_timestep(harness):
    mesh, _, q, bc = S.dam_break_de0(m=10, n=10)
    tau = 0.05
    o = S.setup(OracleDomain(mesh, OracleKernels(gather=True)), "DE1", q, bc)
    o.fixed_flux_timestep = tau
    gen = o.evolve(yieldstep=1e300, finaltime=1e300, max_steps=25)
    next(gen)
    S0 = np.stack([o.stage_c, o.xmom_c, o.ymom_c, o.elev_c], axis=1).copy()
    for _ in gen:
        pass
    assert o.dt_history == [tau] * 25
    L = Layout(mesh, q, bc)
    Sh, _ = _forward(harness, L, S0, tau, 25, [], 1)
    wet = Sh[:, 0] - Sh[:, 3] > 1e-5
    assert 0.1 < wet.mean() < 0.95
    assert np.array_equal(Sh[:, 0], o.stage_c)
    assert np.array_equal(Sh[:, 1], o.xmom_c)
    assert np.array_equal(Sh[:, 2], o.ymom_c)


def _wet_case(m=6, seed=0):
    """Fully wet (depth >= ~2 m) over a smooth bump, moving water, friction
    varying per triangle, transmissive west. Small random noise on bed and
    stage breaks the mesh's mirror symmetries: with exactly equal neighbour
    values the limiter's max / min sits on a kink, where one-sided derivatives
    (the adjoint's) and central differences legitimately differ."""
    p, v, b = rectangular_cross(m, m, 60.0, 60.0)
    mesh = Mesh(p, v, b)
    x, y = mesh.centroid_coordinates.T
    rng = np.random.default_rng(seed)
    z = -4.0 + 1.5 * np.exp(-((x - 30) ** 2 + (y - 25) ** 2) / 200.0) + 1e-3 * rng.random(len(x))
    w = 0.4 * np.exp(-((x - 20) ** 2 + (y - 35) ** 2) / 150.0) + 1e-3 * rng.random(len(x))
    n = 0.03 + 0.01 * rng.random(len(x))
    q = dict(elevation=z, stage=w, friction=n)
    bc = {"left": "transmissive", "right": "reflective", "top": "reflective",
          "bottom": "reflective"}
    return mesh, q, bc


@pytest.mark.parametrize("m,nsteps,obs_every", [(6, 12, 3), (9, 20, 5)])
def test_adjoint_matches_finite_differences(harness, m, nsteps, obs_every):
    mesh, q, bc = _wet_case(m)
    L = Layout(mesh, q, bc)
    N = L.N
    rng = np.random.default_rng(1)
    gauges = rng.choice(N, 6, replace=False).astype(np.int32)
    dt = 0.15    # below the CFL limit of both cases (0.17 for m = 9)
    S0 = L.S.copy()
    x, y = mesh.centroid_coordinates.T       # a smooth initial momentum field
    S0[:, 1] = 0.3 * np.sin(x / 15.0) * np.cos(y / 20.0) + 1e-4 * rng.random(N)
    S0[:, 2] = 0.2 * np.cos(x / 25.0) * np.sin(y / 10.0) + 1e-4 * rng.random(N)
    _, rec = _forward(harness, L, S0, dt, nsteps, gauges, obs_every)
    obs = rec + 0.01 * rng.standard_normal(rec.shape)

    def loss(eta=None, dz=0.0, dS=0.0):
        Lp = Layout(mesh, q, bc)
        if eta is not None:
            Lp.eta[:] = eta
        Sp = S0.copy()
        Sp[:, 3] += dz
        Sp[:, :3] += dS
        _, r = _forward(harness, Lp, Sp, dt, nsteps, gauges, obs_every)
        return float(np.sum((r - obs) ** 2))

    dL = np.ascontiguousarray(2.0 * (rec - obs))
    nbar, zbar, lam = np.zeros(N), np.zeros(N), np.zeros(3 * N)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    harness.harness_adjoint(*L.args(), C.c_double(dt), nsteps, dp(np.ascontiguousarray(S0)),
                            len(gauges), gauges.ctypes.data_as(C.POINTER(C.c_int)), obs_every,
                            dp(dL), dp(nbar), dp(zbar), dp(lam))
    assert np.any(nbar != 0) and np.any(zbar != 0)
    for trial in range(4):
        d = rng.standard_normal(N)
        _agrees(np.dot(nbar, d), lambda e: loss(eta=L.eta + e * d))
        d = rng.standard_normal(N)
        _agrees(np.dot(zbar, d), lambda e: loss(dz=e * d))
        d = rng.standard_normal((N, 3))
        ad = np.dot(lam[:N], d[:, 0]) + np.dot(lam[N:2 * N], d[:, 1]) + np.dot(lam[2 * N:], d[:, 2])
        _agrees(ad, lambda e: loss(dS=e * d))


def smooth_direction(mesh, rng, modes=6):
    """A random direction that is smooth on the mesh's scale: a sum of a few
    random low-order sine modes over the domain's extent. A white-noise bed
    perturbation roughens the bed from one triangle to the next, so every
    small step switches some limiter branch and the central differences
    wander by 1e-5..1e-4 (relative) down to steps of 1e-9, where rounding
    takes over; along smooth directions they settle to 1e-6."""
    x, y = mesh.centroid_coordinates.T
    lx, ly = x.max() - x.min(), y.max() - y.min()
    d = np.zeros(len(x))
    for _ in range(modes):
        kx, ky = rng.integers(1, 4, 2)
        px, py = rng.uniform(0.0, 2.0 * np.pi, 2)
        d += rng.standard_normal() * np.sin(kx * np.pi * x / lx + px) * np.sin(ky * np.pi * y / ly + py)
    return d


def _agrees(ad, f):
    """The central difference at one of the steps 1e-6 .. 1e-9 agrees with ad."""
    diffs = []
    for eps in (1e-6, 1e-7, 1e-8, 1e-9):
        fc = (f(eps) - f(-eps)) / (2 * eps)
        if abs(ad - fc) <= 2e-5 * abs(fc) + 1e-11:
            return eps
        diffs.append(fc)
    raise AssertionError("adjoint %r, central differences %r" % (ad, diffs))


def test_rounding_noise_in_still_water_is_not_amplified(harness):
    """A wave entering still water reaches it first as rounding noise. There
    the limiter's ratio qmin / d of two noise-sized numbers has derivatives
    of size 1/d, and the exact adjoint of a few hundred steps grows to 1e30
    and more. The adjoint holds the limiter's factor at jumps below 1e-10
    (the library's value): the gradient is then bounded and agrees with
    finite differences of the forward (which the freeze does not change)."""
    import bench
    harness.harness_set_freeze.argtypes = [C.c_double]
    mesh, q, bc, gauges = bench.build_c4(m=16, length=5000.0)
    L = Layout(mesh, q, bc)
    N = L.N
    S0 = np.ascontiguousarray(np.stack([q["stage"], 0 * q["stage"], 0 * q["stage"],
                                        q["elevation"]], axis=1))
    g = np.ascontiguousarray(gauges[:16] % N, np.int32)
    nsteps, dt = 300, 0.1
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    _, rec = _forward(harness, L, S0, dt, nsteps, g, 10)
    obs = rec + 1e-3 * np.random.default_rng(0).standard_normal(rec.shape)
    dL = np.ascontiguousarray(2.0 * (rec - obs))

    def grads(freeze):
        harness.harness_set_freeze(freeze)
        nbar, zbar, lam = np.zeros(N), np.zeros(N), np.zeros(3 * N)
        harness.harness_adjoint(*L.args(), C.c_double(dt), nsteps, dp(S0), len(g),
                                g.ctypes.data_as(C.POINTER(C.c_int)), 10, dp(dL), dp(nbar),
                                dp(zbar), dp(lam))
        return nbar, zbar
    try:
        _, z_exact = grads(0.0)
        nbar, zbar = grads(1.0e-10)
    finally:
        harness.harness_set_freeze(1.0e-10)
    assert np.abs(z_exact).max() > 1e10
    assert np.abs(zbar).max() < 1e-2

    def loss(eta=None, dz=0.0):
        Lp = Layout(mesh, q, bc)
        if eta is not None:
            Lp.eta[:] = eta
        Sp = S0.copy()
        Sp[:, 3] += dz
        _, r = _forward(harness, Lp, Sp, dt, nsteps, g, 10)
        return float(np.sum((r - obs) ** 2))
    rng = np.random.default_rng(5)
    d = rng.standard_normal(N)
    fd = (loss(eta=L.eta + 1e-6 * d) - loss(eta=L.eta - 1e-6 * d)) / 2e-6
    assert np.dot(nbar, d) == pytest.approx(fd, rel=1e-4)
    d = rng.standard_normal(N)
    fd = (loss(dz=1e-6 * d) - loss(dz=-1e-6 * d)) / 2e-6
    assert np.dot(zbar, d) == pytest.approx(fd, rel=0.05)


ASAN_SCRIPT = r'''
import ctypes as C, sys
import numpy as np
sys.path.insert(0, sys.argv[2])
from tests.test_adjoint_cpu import Layout, _forward, _wet_case
lib = C.CDLL(sys.argv[1])
dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
mesh, q, bc = _wet_case(16)
L = Layout(mesh, q, bc)
N = L.N
rng = np.random.default_rng(4)
x, y = mesh.centroid_coordinates.T
S0 = L.S.copy()
S0[:, 1] = 0.3 * np.sin(x / 15.0) * np.cos(y / 20.0)
S0[:, 2] = 0.2 * np.cos(x / 25.0) * np.sin(y / 10.0)
g = rng.choice(N, 12, replace=False).astype(np.int32)
for dt, nsteps, expect_finite in ((0.08, 60, True), (0.6, 40, False)):
    S, rec = _forward(lib, L, S0, dt, nsteps, g, 10)
    assert np.all(np.isfinite(S)) == expect_finite, (dt, "finite" if expect_finite else "blow-up")
    dL = np.ascontiguousarray(np.nan_to_num(rec))
    nbar, zbar, lam = np.zeros(N), np.zeros(N), np.zeros(3 * N)
    lib.harness_adjoint(*L.args(), C.c_double(dt), nsteps, dp(np.ascontiguousarray(S0)), len(g),
                        g.ctypes.data_as(C.POINTER(C.c_int)), 10, dp(dL), dp(nbar), dp(zbar),
                        dp(lam))
xs = np.array([np.nan, np.inf, -np.inf, 0.0, -0.0, -1.0, 5e-324, 1e-310, 1e300, 1e-300, 1.0])
out = np.empty_like(xs)
lib.pow73_many(len(xs), dp(xs), dp(out))
print("asan run ok")
'''

POW_SHIM = r'''
#include "de_pow.h"
extern "C" void pow73_many(long n, const double *x, double *y) {
    for (long i = 0; i < n; i++) y[i] = de_pow_7_3(x[i]);
}
'''


def test_adjoint_cell_math_in_bounds_under_asan():
    """The adjoint's cell math (the device's, compiled for the host) under
    AddressSanitizer: the 16x16 gradient case at dt 0.08 (below its CFL limit
    of 0.095) stays finite, a run far above the limit (dt 0.6) blows up to
    NaN, and both the forward and the adjoint sweep of either stay in bounds,
    as does pow on NaN, infinite, zero, negative and subnormal depths."""
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                          text=True).stdout.strip()
    if not os.path.isabs(asan) or not os.path.exists(asan):
        pytest.skip("no AddressSanitizer runtime")
    tmp = tempfile.mkdtemp()
    shim = os.path.join(tmp, "shim.cpp")
    open(shim, "w").write(POW_SHIM)
    so = os.path.join(tmp, "libharness_asan.so")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-ffp-contract=off", "-shared", "-fPIC",
                    "-fsanitize=address", "-fno-omit-frame-pointer",
                    "-I" + os.path.join(ROOT, "anuga_core_amd", "csrc"), "-o", so,
                    os.path.join(ROOT, "tests", "adjoint_harness.cpp"), shim], check=True)
    script = os.path.join(tmp, "run.py")
    open(script, "w").write(ASAN_SCRIPT)
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([sys.executable, script, so, ROOT], env=env, capture_output=True,
                       text=True, timeout=600)
    assert "AddressSanitizer" not in r.stderr, r.stderr[-3000:]
    assert r.returncode == 0 and "asan run ok" in r.stdout, r.stderr[-3000:]


@pytest.mark.parametrize("case", ["wet", "dam_break", "dam_break_frozen", "corners"])
def test_reverse_extrapolation_product_equals_dual(harness, case):
    """The extrapolation's hand-written reverse sweep (adj_extrap_vjp) gives
    the forward-mode dual product's adjoints per stencil slot to rounding, on
    moving wet water, on a wet/dry dam break (one-boundary triangles, dry
    neighbours, clamped cells), with and without the frozen limiter factor,
    and on a `rectangular` mesh whose corner triangles have two boundary edges
    (the one-neighbour gradient)."""
    if case == "wet":
        mesh, q, bc = _wet_case(9)
        L = Layout(mesh, q, bc)
        Sx, _ = _forward(harness, L, L.S, 0.1, 8, [], 1)
    elif case == "corners":
        mesh = Mesh(*rectangular(8, 6, 40.0, 30.0))
        x, y = mesh.centroid_coordinates.T
        z = -0.02 * x + 0.01 * np.sin(y)
        q = dict(elevation=z, stage=np.where(x < 20, 1.0, z),
                 friction=np.full(mesh.number_of_triangles, 0.03))
        bc = {t: "reflective" for t in ("left", "right", "top", "bottom")}
        L = Layout(mesh, q, bc)
        Sx, _ = _forward(harness, L, L.S, 0.02, 20, [], 1)
        assert (mesh.number_of_boundaries == 2).any()
    else:
        mesh, _, q, bc = S.dam_break_de0(m=12, n=10)
        L = Layout(mesh, q, bc)
        Sx, _ = _forward(harness, L, L.S, 0.02, 15, [], 1)
    harness.harness_set_freeze(C.c_double(1e-9 if case == "dam_break_frozen" else 0.0))
    N = L.N
    wt = np.random.default_rng(5).standard_normal(12 * N)
    Xd, Xr = np.zeros(16 * N), np.zeros(16 * N)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    S_ = np.ascontiguousarray(Sx)
    harness.harness_extrap_products(*L.args(), dp(S_), dp(wt), dp(Xd), dp(Xr))
    harness.harness_set_freeze(C.c_double(0.0))
    scale = np.abs(Xd).max()
    assert scale > 0
    assert np.max(np.abs(Xr - Xd)) <= 1e-11 * scale, np.max(np.abs(Xr - Xd)) / scale
    # every slot that carries a value in one carries it in the other
    assert np.array_equal(Xd != 0.0, Xr != 0.0) or np.max(np.abs(Xr - Xd)) == 0.0


@pytest.mark.parametrize("case", ["wet", "dam_break", "boundaries"])
def test_reverse_flux_product_equals_dual(harness, case):
    """The flux's hand-written reverse sweep (adj_edge_vjp) gives the
    forward-mode dual product's adjoints of the edge values, the centroid
    state and the bed to rounding: on moving wet water, on a wet/dry dam break
    (dry edges, the zero-depth early return, clamped wave speeds), and on a
    mesh whose boundary edges are reflective, transmissive, Dirichlet and
    centroid-transmissive."""
    if case == "wet":
        mesh, q, bc = _wet_case(9)
        steps, dt = 8, 0.1
    elif case == "boundaries":
        mesh = Mesh(*rectangular(8, 6, 40.0, 30.0))
        x, y = mesh.centroid_coordinates.T
        z = -0.02 * x + 0.01 * np.sin(y)
        q = dict(elevation=z, stage=np.where(x < 20, 1.0, z),
                 xmomentum=np.where(x < 20, 0.1, 0.0),
                 friction=np.full(mesh.number_of_triangles, 0.03))
        bc = {"left": ("dirichlet", [0.8, 0.05, 0.0]), "right": "transmissive",
              "top": "reflective", "bottom": "transmissive_centroid"}
        steps, dt = 20, 0.02
    else:
        mesh, _, q, bc = S.dam_break_de0(m=12, n=10)
        steps, dt = 15, 0.02
    L = Layout(mesh, q, bc)
    Sx, _ = _forward(harness, L, L.S, dt, steps, [], 1)
    N = L.N
    wt = np.random.default_rng(6).standard_normal(3 * N)
    Od, Or = np.zeros(16 * N), np.zeros(16 * N)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    S_ = np.ascontiguousarray(Sx)
    harness.harness_flux_products(*L.args(), dp(S_), dp(wt), dp(Od), dp(Or))
    # lamE, lamS, zbar, each to rounding of its own scale
    for a, b in ((0, 12 * N), (12 * N, 15 * N), (15 * N, 16 * N)):
        d, r = Od[a:b], Or[a:b]
        scale = np.abs(d).max()
        assert scale > 0
        assert np.max(np.abs(r - d)) <= 1e-11 * scale, np.max(np.abs(r - d)) / scale

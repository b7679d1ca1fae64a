"""Parity of the HIP DE step with the CPU oracle (runs on an MI355X).

Each case evolves the same seeded scenario through the device Domain (the C ABI
of libanuga_hip.so via ctypes) and through the oracle (the C restatement of
the reference's FP64 path, itself pinned bit for bit to the reference's
compiled C), then compares:
  * the timestep sequence: identical;
  * stage / xmomentum / ymomentum centroid values: bit for bit (FP64 path;
    the oracle calls the C library's pow as the reference does, and the
    device's pow returns the same bits);
  * diagnostics summed in a different order (volume, boundary flux integral):
    relative tolerance 1e-12.
"""

import numpy as np
import pytest

from anuga_core_amd.domain import Domain
from oracle.kernels import OracleKernels
from oracle.oracle_domain import OracleDomain
from tests import scenarios as S

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _pair(builder, betas=None, reorder=True, **kw):
    mesh, alg, q, bc = builder(**kw)
    o = S.setup(OracleDomain(mesh, OracleKernels(gather=True)), alg, q, bc, betas)
    d = S.setup(Domain(mesh=mesh, reorder=reorder), alg, q, bc, betas)
    return o, d


def _compare(o, d, label):
    for name, oname in (("stage", "stage_c"), ("xmomentum", "xmom_c"), ("ymomentum", "ymom_c")):
        a = d.get_quantity(name).centroid_values
        b = getattr(o, oname)
        err = S.rel_maxnorm(a, b)
        print("%s %s rel max-norm %.3e bitwise=%s" % (label, name, err, np.array_equal(a, b)),
              flush=True)
        assert np.array_equal(a, b), (label, name, err)


def test_flat_bed_golden():
    """The reference's test_second_order_flat_bed_onestep on the device."""
    print("case flat_bed_golden", flush=True)
    o, d = _pair(S.flat_bed_dirichlet, betas=0.9)
    for _ in d.evolve(yieldstep=0.05, finaltime=0.05):
        pass
    for _ in o.evolve(yieldstep=0.05, finaltime=0.05):
        pass
    assert np.allclose(d.recorded_min_timestep, 0.03571428571428564)
    w = d.get_quantity("stage").centroid_values
    assert np.allclose(w[:12], [0.00117244, 0.025897, 0.00200148, 0.025897, 0.00200148,
                                0.025897, 0.00200148, 0.025897, 0.00200148, 0.025897,
                                0.00200148, 0.02672604])
    assert np.allclose(d.get_quantity("xmomentum").centroid_values[:12],
                       [0.000189, 0.0042] * 6)
    d.distribute_to_vertices_and_edges()
    assert np.allclose(d.get_quantity("stage").vertex_values[:12, 0],
                       [-0.00104301, 0.025897, 0.00020015, 0.025897, 0.00020015, 0.025897,
                        0.00020015, 0.025897, 0.00020015, 0.025897, 0.00011501, 0.02672604])
    _compare(o, d, "flat_bed")


@pytest.mark.parametrize("reorder", [True, "rcm", False])
def test_de0_dam_break_vs_oracle(reorder):
    print("case de0_dam_break reorder=%s" % reorder, flush=True)
    o, d = _pair(S.dam_break_de0, reorder=reorder, m=40, n=40)
    for _ in d.evolve(yieldstep=1.0, finaltime=4.0):
        pass
    for _ in o.evolve(yieldstep=1.0, finaltime=4.0):
        pass
    assert d.dt_history == o.dt_history, "timestep sequences differ"
    _compare(o, d, "de0")
    assert abs(d.get_water_volume() - o.get_water_volume()) <= TOL * o.get_water_volume()


@pytest.mark.parametrize("reorder", [True, "rcm"])
def test_de1_hills_vs_oracle(reorder):
    """Bisection tiles (fused extrapolation + flux) and RCM (separate kernels)."""
    print("case de1_hills reorder=%s" % reorder, flush=True)
    o, d = _pair(S.hills_de1, reorder=reorder, m=30, n=30)
    for _ in d.evolve(yieldstep=5.0, finaltime=40.0):
        pass
    for _ in o.evolve(yieldstep=5.0, finaltime=40.0):
        pass
    print("steps", len(d.dt_history), len(o.dt_history), flush=True)
    assert d.dt_history == o.dt_history, "timestep sequences differ"
    _compare(o, d, "de1")
    bo, bd = o.boundary_flux_integral, d.boundary_flux_integral
    assert abs(bd - bo) <= TOL * max(abs(bo), 1.0)


def test_graded_mesh_mixed_same_bits_in_every_order():
    """Mixed precision on the unstructured mesh: the fused kernels (tiles) and
    the separate ones (RCM) give the same state bits."""
    print("case graded mesh mixed orders", flush=True)
    runs = []
    for reorder in (True, "rcm"):
        mesh, alg, q, bc = S.hills_graded()
        d = S.setup(Domain(mesh=mesh, precision="mixed", reorder=reorder), alg, q, bc)
        for _ in d.evolve(yieldstep=5.0, finaltime=10.0):
            pass
        runs.append((list(d.dt_history), d.get_quantity("stage").centroid_values,
                     d.get_quantity("xmomentum").centroid_values))
    assert runs[0][0] == runs[1][0]
    assert np.array_equal(runs[0][1], runs[1][1]) and np.array_equal(runs[0][2], runs[1][2])


@pytest.mark.parametrize("reorder", [True, "rcm"])
def test_graded_unstructured_mesh_vs_oracle(reorder):
    """The unstructured (Delaunay, two densities, 2-boundary-edge corner
    triangles) mesh of the large-mesh variant, scaled down: bit for bit."""
    print("case graded mesh reorder=%s" % reorder, flush=True)
    o, d = _pair(S.hills_graded, reorder=reorder)
    for _ in d.evolve(yieldstep=5.0, finaltime=20.0):
        pass
    for _ in o.evolve(yieldstep=5.0, finaltime=20.0):
        pass
    print("steps", len(d.dt_history), flush=True)
    assert d.dt_history == o.dt_history, "timestep sequences differ"
    for name, oname in (("stage", "stage_c"), ("xmomentum", "xmom_c"), ("ymomentum", "ymom_c")):
        assert np.array_equal(d.get_quantity(name).centroid_values, getattr(o, oname)), name


def test_edge_values_after_distribute():
    print("case edge_values", flush=True)
    o, d = _pair(S.hills_de1, m=20, n=20)
    o.distribute_to_vertices_and_edges()
    d.distribute_to_vertices_and_edges()
    N = o.N
    for name, oname in (("stage", "stage_e"), ("height", "height_e"), ("xmomentum", "xmom_e"),
                        ("ymomentum", "ymom_e"), ("elevation", "elev_e")):
        a = d.get_quantity(name).edge_values
        b = getattr(o, oname).reshape(N, 3)
        assert np.array_equal(a, b), name
    for name, oname in (("stage", "stage_c"), ("xmomentum", "xmom_c"), ("ymomentum", "ymom_c")):
        assert np.array_equal(d.get_quantity(name).centroid_values, getattr(o, oname)), name


def test_edge_values_after_fused_steps():
    """The fused kernels keep edge values on chip: asked for after device
    steps, the library extrapolates the current state, which equals what
    distribute_to_vertices_and_edges then computes, and what the separate
    kernels (RCM order) give for the same state."""
    print("case edge_values after fused steps", flush=True)
    runs = []
    for reorder in (True, "rcm"):
        mesh, alg, q, bc = S.hills_de1(m=20, n=20)
        d = S.setup(Domain(mesh=mesh, reorder=reorder), alg, q, bc)
        d.evolve_steps(25)
        before = {n: d.get_quantity(n).edge_values for n in ("stage", "height", "xmomentum")}
        d.distribute_to_vertices_and_edges()
        after = {n: d.get_quantity(n).edge_values for n in ("stage", "height", "xmomentum")}
        runs.append((before, after))
    (b_t, a_t), (b_r, a_r) = runs
    for n in a_t:
        assert np.array_equal(b_t[n], a_t[n]), n
        assert np.array_equal(a_t[n], a_r[n]), n


def test_run_twice_bitwise():
    """Fixed-order gathers and reductions: two runs are bit-identical."""
    print("case determinism", flush=True)
    res = []
    for _ in range(2):
        mesh, alg, q, bc = S.hills_de1(m=24, n=24)
        d = S.setup(Domain(mesh=mesh), alg, q, bc)
        d.evolve_steps(50)
        res.append((d.get_quantity("stage").centroid_values,
                    d.get_quantity("xmomentum").centroid_values, list(d.dt_history),
                    d.boundary_flux_integral))
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])
    assert res[0][2] == res[1][2]
    assert res[0][3] == res[1][3]


def test_volume_conservation_reflective():
    """Closed basin: volume is constant to rounding (test_conservation_* pattern)."""
    print("case conservation", flush=True)
    mesh, alg, q, bc = S.hills_de1(m=24, n=24)
    bc = {t: "reflective" for t in bc}
    d = S.setup(Domain(mesh=mesh), alg, q, bc)
    d.distribute_to_vertices_and_edges()
    v0 = d.get_water_volume()
    d.evolve_steps(200)
    d.distribute_to_vertices_and_edges()
    v1 = d.get_water_volume()
    assert abs(v1 - v0) <= 1e-9 * v0, (v0, v1)


@pytest.mark.parametrize("builder,kw,ys,ft", [
    (S.dam_break_de0, dict(m=30, n=30), 0.5, 3.0),
    (S.hills_de1, dict(m=24, n=24), 4.0, 30.0),
])
def test_device_timestep_matches_host_timestep(builder, kw, ys, ft):
    """Runs of device-controlled steps (update_timestep on the device, one
    captured HIP graph per step replayed, host syncs only at yields) equal the
    host-controlled per-substep sequence bit for bit, at every yield."""
    print("case device_vs_host_timestep %s" % builder.__name__, flush=True)
    runs = []
    for device in (False, True):
        mesh, alg, q, bc = builder(**kw)
        d = S.setup(Domain(mesh=mesh), alg, q, bc)
        d.device_timestep = device
        snaps = []
        for t in d.evolve(yieldstep=ys, finaltime=ft):
            snaps.append((t, d.get_quantity("stage").centroid_values,
                          d.get_quantity("xmomentum").centroid_values,
                          d.recorded_min_timestep, d.recorded_max_timestep))
        runs.append((snaps, list(d.dt_history), d.boundary_flux_integral))
    (sa, ha, ia), (sb, hb, ib) = runs
    assert ha == hb
    assert len(sa) == len(sb)
    for a, b in zip(sa, sb):
        assert a[0] == b[0] and a[3] == b[3] and a[4] == b[4]
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    assert ia == ib


def test_graph_replay_matches_eager(monkeypatch):
    """The same device-controlled steps with and without HIP graph replay,
    interleaved with eager host work (reads, volume reductions) between runs."""
    print("case graph_vs_eager", flush=True)
    res = []
    for graph in ("1", "0"):
        monkeypatch.setenv("ANUGA_HIP_GRAPH", graph)
        mesh, alg, q, bc = S.hills_de1(m=20, n=20, seed=4)
        d = S.setup(Domain(mesh=mesh), alg, q, bc)
        vals = []
        for _ in range(4):
            d.evolve_steps(13)
            vals.append((d.get_quantity("stage").centroid_values, d.get_water_volume(),
                         list(d.dt_history)))
        res.append(vals)
    for a, b in zip(*res):
        assert np.array_equal(a[0], b[0]) and a[1] == b[1] and a[2] == b[2]


def _error_report(a, b):
    d = np.abs(a - b)
    return float(d.sum() / max(np.abs(b).sum(), 1e-300)), float(d.max())


@pytest.mark.parametrize("builder,kw,ft", [(S.hills_de1, dict(m=30, n=30), 30.0),
                                           (S.dam_break_de0, dict(m=30, n=30), 3.0)])
def test_mixed_precision_error_vs_oracle(builder, kw, ft):
    """FP32 extrapolation / reconstruction / flux with FP64 state, update, CFL
    and thresholds: stage relative L1 error <= 1e-5 against the FP64 oracle;
    volume conserved to FP64 rounding of the same accounting."""
    print("case mixed %s" % builder.__name__, flush=True)
    mesh, alg, q, bc = builder(**kw)
    o = S.setup(OracleDomain(mesh, OracleKernels(gather=True)), alg, q, bc)
    d = S.setup(Domain(mesh=mesh, precision="mixed"), alg, q, bc)
    for _ in o.evolve(yieldstep=ft, finaltime=ft):
        pass
    for _ in d.evolve(yieldstep=ft, finaltime=ft):
        pass
    l1w, mxw = _error_report(d.get_quantity("stage").centroid_values, o.stage_c)
    l1u, mxu = _error_report(d.get_quantity("xmomentum").centroid_values, o.xmom_c)
    print("mixed stage L1 %.3e max %.3e | xmom L1 %.3e max %.3e | steps %d vs %d"
          % (l1w, mxw, l1u, mxu, len(d.dt_history), len(o.dt_history)), flush=True)
    h = np.maximum(o.stage_c - o.elev_c, 0.0)
    l1h = float(np.abs(d.get_quantity("stage").centroid_values - o.stage_c).sum() / h.sum())
    print("mixed stage L1 relative to the water column %.3e" % l1h, flush=True)
    assert l1h <= 1e-5
    v0 = float(np.sum(mesh.areas * np.maximum(q["stage"] - q["elevation"], 0)))
    drift = (d.get_water_volume() - v0 - d.boundary_flux_integral) / v0
    print("mixed volume drift %.3e" % drift, flush=True)
    assert abs(drift) < 1e-10


def _with_algorithm(builder, alg):
    def b(**kw):
        mesh, _, q, bc = builder(**kw)
        return mesh, alg, q, bc
    b.__name__ = "%s_%s" % (builder.__name__, alg)
    return b


@pytest.mark.parametrize("alg,builder,kw,ys,ft", [
    ("DE2", S.hills_de1, dict(m=24, n=24), 5.0, 25.0),
    ("DE2", S.dam_break_de0, dict(m=30, n=30), 1.0, 3.0),
    ("DE1_7", S.hills_de1, dict(m=24, n=24), 5.0, 25.0),
    ("DE0_7", S.dam_break_de0, dict(m=30, n=30), 1.0, 3.0),
])
def test_other_de_algorithms_vs_oracle(alg, builder, kw, ys, ft):
    """DE2 (rk3) and the DE0_7 / DE1_7 presets (beta_dry 0.1) against the oracle
    (shallow_water_domain.py:709-895), device-controlled steps."""
    print("case %s %s" % (alg, builder.__name__), flush=True)
    o, d = _pair(_with_algorithm(builder, alg), **kw)
    for _ in d.evolve(yieldstep=ys, finaltime=ft):
        pass
    for _ in o.evolve(yieldstep=ys, finaltime=ft):
        pass
    assert d.dt_history == o.dt_history, "timestep sequences differ"
    _compare(o, d, alg)
    bo, bd = o.boundary_flux_integral, d.boundary_flux_integral
    assert abs(bd - bo) <= TOL * max(abs(bo), 1.0)


def test_de2_device_timestep_matches_host_timestep():
    test_device_timestep_matches_host_timestep(_with_algorithm(S.hills_de1, "DE2"),
                                               dict(m=20, n=20), 4.0, 20.0)


from tests import golden_cases as GC  # noqa: E402


@pytest.mark.parametrize("case", GC.CASES, ids=[c.__name__[5:] for c in GC.CASES])
def test_golden_multi_yield_on_device(case):
    """The reference's multi-yield golden trajectories on the device, and the
    device bit for bit with the oracle on the same case."""
    print("case golden %s" % case.__name__, flush=True)
    d = case(lambda mesh: Domain(mesh=mesh))
    o = case(lambda mesh: OracleDomain(mesh, OracleKernels(gather=True), "DE0"))
    assert d.dt_history == o.dt_history
    _compare(o, d, case.__name__)

"""Benchmark: M triangle-steps/s of the DE1 step on a 16M-triangle mesh.

    python bench.py [--gpus 1] [--steps 1000] [--warmup 20] [--precision fp64|mixed]
                    [--m 2000] [--cpu-steps 20] [--no-cpu] [--no-compare]

Workload (one GPU): rectangular_cross(2000, 2000) over 20 km x 20 km, 16,000,000
triangles, flow algorithm DE1 (rk2, FP64). Bed: -50 m plus 64 Gaussian hills
(amplitude U(5,60) m, sigma U(200,1500) m, centres uniform;
numpy.random.default_rng(0)). Stage: a 3 m Gaussian hump of sigma 1 km at the
centre, clipped to the bed (dry where the bed is higher). Manning n = 0.025.
Boundaries: transmissive west, reflective elsewhere. Synthetic data.

A step is one full DE1 step (two extrapolate + flux substeps, friction,
update, saxpy, CFL reduction and host timestep control); the timed region runs
exactly --steps steps with the state already resident in HBM, bracketed by
device synchronisation. Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per triangle, each input of the unit read once and each
# output written once (neighbour gathers are re-reads and do not count).
#   Whole DE1 step (the reference's data model, int32 indices): 2 x 232 B per
#   Euler substep + 24 B for reading the step's initial state in the final
#   saxpy = 488 B (FP64); 344 B with FP32 geometry (mixed).
#   Per kernel (this implementation's split of a substep), short name ->
#   (kernel, bytes):
KERNEL_BYTES = {
    "fp64": {
        # w,uh,vh,z (32) + 3 neighbours (12) + centroid (16) + edge offsets (48)
        #   -> w,h,uh,vh on 3 edges (96) + flag (1)
        "extrapolate": ("de_extrapolate", 205),
        # edge values (96) + w,z (16) + flag (1) + neighbours (12) + normals (48)
        #   + lengths (24) + area, radius (16) -> explicit updates (24)
        "flux0": ("de_flux<0>", 237),
        # w,uh,vh,z (32) + n (8) + flag (1) + explicit updates (24) -> w,uh,vh (24)
        "update": ("de_update", 89),
        # edge values (96) + w,uh,vh,z (32) + flag (1) + neighbours (12)
        #   + normals (48) + lengths (24) + area (8) + n (8) + step-initial
        #   w,uh,vh (24) -> w,uh,vh (24)
        "flux1": ("de_flux<1>", 277),
        # the tiled fused kernels (compact tiles): extrapolation inputs w,uh,vh,z
        #   (32) + neighbours (12) + centroid (16) + edge offsets (48) + normals
        #   (48) + lengths (24) + area, radius (16) -> flag (1) + explicit (24)
        "xflux0": ("de_xflux<0>", 221),
        # extrapolation inputs (108) + normals (48) + lengths (24) + area (8)
        #   + n (8) + step-initial w,uh,vh (24) -> flag (1) + w,uh,vh (24)
        "xflux1": ("de_xflux<1>", 245),
    },
    "mixed": {
        # w,uh,vh,z (32) + neighbours (12) + FP32 edge and neighbour offsets (48)
        #   -> FP32 jumps dw,dh,uh,vh on 3 edges (48) + FP64 depth (8) + flag (1)
        "extrapolate": ("de_extrapolate", 149),
        # edge record (56) + w,z (16) + flag (1) + neighbours (12) + FP32 normals,
        #   lengths, radius, neighbour bed offsets (52) + area (8) -> explicit (24)
        "flux0": ("de_flux_mixed<0>", 169),
        "update": ("de_update", 89),
        # edge record (56) + w,uh,vh,z (32) + flag (1) + neighbours (12) + FP32
        #   geometry without radius (48) + area (8) + n (8) + step-initial
        #   w,uh,vh (24) -> w,uh,vh (24)
        "flux1": ("de_flux_mixed<1>", 213),
        # fused over tiles: w,uh,vh,z (32) + neighbours (12) + FP32 edge and
        #   neighbour offsets (48) + FP32 normals, lengths, radius, neighbour bed
        #   offsets (52) + area (8) -> flag (1) + explicit updates (24)
        "xflux0": ("de_xflux_mixed<0>", 177),
        # extrapolation inputs (92) + FP32 geometry without radius (48) + area (8)
        #   + n (8) + step-initial w,uh,vh (24) -> flag (1) + w,uh,vh (24)
        "xflux1": ("de_xflux_mixed<1>", 205),
    },
}
LAUNCHES_PER_STEP = {"extrapolate": 2, "flux0": 1, "update": 1, "flux1": 1, "xflux0": 1,
                     "xflux1": 1}
STEP_BYTES = {"fp64": 488, "mixed": 344}
PROFILE = os.path.join(ROOT, "profiles", "de_kernels_rectangular_cross2000_DE1_%s.txt")
PROFILE_GRADED = os.path.join(ROOT, "profiles", "de_kernels_graded_square4M_DE1_%s.txt")


def measured_traffic(precision, kernel, m, kind="rectangular_cross"):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary of
    this workload (FETCH_SIZE and WRITE_SIZE, scaled by a 1 GiB copy), or None
    when there is no summary for this mesh and kernel."""
    if kind == "graded":
        path = PROFILE_GRADED % precision
    elif m == 2000:
        path = PROFILE % precision
    else:
        return None
    if not os.path.exists(path):
        return None
    header = None
    with open(path) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            cols = line.split()
            if header is None:
                header = cols
                continue
            if cols[0] == kernel and "hbm_bytes" in header:
                return float(cols[header.index("hbm_bytes")])
    return None


def kernel_label(kname):
    """The fused FP64 kernels run as de_xflux_e (each edge's flux evaluated
    once) unless ANUGA_HIP_EDGE_ONCE=0; the meshes here mirror exactly."""
    if kname.startswith("de_xflux<") and os.environ.get("ANUGA_HIP_EDGE_ONCE", "1") != "0":
        return kname.replace("de_xflux<", "de_xflux_e<")
    return kname


def _reorder(args):
    return {"tiles": True, "rcm": "rcm", "none": False}["none" if args.no_reorder else args.order]


def build_c1(m=2000, length=20000.0, nhills=64, seed=0, kind="rectangular_cross"):
    """Mesh and initial condition of the 16M-triangle DE1 workload; with
    kind='graded', the unstructured variant: about 4M triangles of 50 m^2
    inside the central 13 km square and 400 m^2 outside (graded_square,
    seed 0), the same bed, stage, friction and boundaries."""
    from anuga_core_amd.mesh import Mesh, graded_square, rectangular_cross
    if kind == "graded":
        f0, f1 = 0.175 * length, 0.825 * length
        p, v, b = graded_square(length, 50.0, 400.0, (f0, f0, f1, f1), seed=0)
    else:
        p, v, b = rectangular_cross(m, m, length, length)
    mesh = Mesh(p, v, b)
    del p, v
    rng = np.random.default_rng(seed)
    amp = rng.uniform(5.0, 60.0, nhills)
    sig = rng.uniform(200.0, 1500.0, nhills) * (length / 20000.0)
    cx = rng.uniform(0.0, length, nhills)
    cy = rng.uniform(0.0, length, nhills)
    x = mesh.centroid_coordinates[:, 0]
    y = mesh.centroid_coordinates[:, 1]
    z = np.full(x.shape, -50.0)
    for j in range(nhills):
        z += amp[j] * np.exp(-((x - cx[j]) ** 2 + (y - cy[j]) ** 2) / (2.0 * sig[j] ** 2))
    s = 1000.0 * (length / 20000.0)
    hump = 3.0 * np.exp(-((x - length / 2) ** 2 + (y - length / 2) ** 2) / (2.0 * s * s))
    w = np.maximum(hump, z)
    q = dict(elevation=z, stage=w, friction=np.full(x.shape, 0.025))
    bc = {"left": "transmissive", "right": "reflective", "top": "reflective",
          "bottom": "reflective"}
    return mesh, q, bc


def _boundaries(domain, bc):
    from anuga_core_amd import Reflective_boundary, Transmissive_boundary
    make = {"reflective": Reflective_boundary, "transmissive": Transmissive_boundary}
    return {t: make[k](domain) for t, k in bc.items()}


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo, as lscpu reports it)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(mesh, q, bc, steps, threads=None):
    """The oracle (C restatement of the reference's FP64 path, OpenMP per-
    triangle loops, libm pow as the reference) on this host, timed over a
    bounded number of steps of the same workload; `threads` OpenMP threads
    (default: the OpenMP default, OMP_NUM_THREADS). Returns (steps, seconds,
    threads used)."""
    import ctypes
    from oracle.kernels import OracleKernels
    from oracle.oracle_domain import OracleDomain
    gomp = ctypes.CDLL("libgomp.so.1")   # the OpenMP runtime the oracle links
    prev = gomp.omp_get_max_threads()
    if threads:
        gomp.omp_set_num_threads(int(threads))
    used = gomp.omp_get_max_threads()
    try:
        o = OracleDomain(mesh, OracleKernels(gather=True), "DE1")
        for name, val in q.items():
            o.set_centroid_values(name, val)
        o.set_boundary({t: type("B", (), {"kind": k})() for t, k in bc.items()})
        gen = o.evolve(yieldstep=1.0e300, finaltime=1.0e300, max_steps=steps)
        next(gen)   # the initial yield comes before any step
        t0 = time.perf_counter()
        for _ in gen:
            pass
        dt = time.perf_counter() - t0
    finally:
        gomp.omp_set_num_threads(prev)
    return len(o.dt_history), dt, used


def run_timed(precision, mesh, q, bc, args):
    """Set up the device domain, run the warm-up, the timed steps and the
    eager per-kernel timing pass. Returns (domain, result dict)."""
    from anuga_core_amd.domain import Domain
    t0 = time.perf_counter()
    d = Domain(mesh=mesh, precision=precision, reorder=_reorder(args))
    d.set_flow_algorithm("DE1")
    for name, val in q.items():
        d.set_quantity(name, val, location="centroids")
    d.set_boundary(_boundaries(d, bc))
    d.record_dt_history = False
    print("%s domain ready in %.1fs" % (precision, time.perf_counter() - t0),
          file=sys.stderr, flush=True)
    d.distribute_to_vertices_and_edges()
    v0 = d.get_water_volume()
    d.evolve_steps(args.warmup)
    d.synchronize()
    # timed region: device-controlled steps, one HIP graph per step
    t1 = time.perf_counter()
    d.evolve_steps(args.steps)
    d.synchronize()
    elapsed = time.perf_counter() - t1
    print("%s: %d steps in %.3fs" % (precision, args.steps, elapsed), file=sys.stderr, flush=True)
    # per-kernel durations: HIP events around every launch on the library's
    # stream, over an eager (non-graph) run of further steps
    d.reset_kernel_timing(True)
    d.evolve_steps(args.timing_steps)
    d.synchronize()
    kt = {n: d.kernel_timing(n) for n in KERNEL_BYTES[precision]}
    d.reset_kernel_timing(False)
    d.distribute_to_vertices_and_edges()
    v1 = d.get_water_volume()
    N = mesh.number_of_triangles
    kernels = {}
    for short, (ms, n) in kt.items():
        if n == 0:
            continue   # not run in this configuration (fused or separate kernels)
        kname, nbytes = KERNEL_BYTES[precision][short]
        kernels[kernel_label(kname)] = {"avg_ms": round(ms, 5), "launches": n,
                          "alg_bytes_per_triangle": nbytes,
                          "alg_GBs": round(nbytes * N / (ms * 1e-3) / 1e9, 1) if ms > 0 else None}
    ms_per_step = 1e3 * elapsed / args.steps
    return d, {
        "value": N * args.steps / elapsed / 1e6,
        "ms_per_step": ms_per_step,
        "kernels": kernels,
        "volume_drift": (v1 - v0 - d.boundary_flux_integral) / v0,
        "steps_total": args.warmup + args.steps + args.timing_steps,
    }


def error_report(mixed, ref):
    """Mixed-precision state against the FP64 device state after the same
    number of steps: L1 (sum |diff| / sum |ref|) and max |diff| per quantity,
    and the stage L1 relative to the water column (sum |diff| / sum h)."""
    out = {}
    for name in ("stage", "xmomentum", "ymomentum"):
        a = mixed.get_quantity(name).centroid_values
        b = ref.get_quantity(name).centroid_values
        d = np.abs(a - b)
        out[name] = {"L1": float(d.sum() / np.abs(b).sum()), "max": float(d.max())}
        if name == "stage":
            h = np.maximum(b - ref.get_quantity("elevation").centroid_values, 0.0)
            out["stage_L1_over_water_column"] = float(d.sum() / h.sum())
    out["time_difference_s"] = mixed.get_time() - ref.get_time()
    return out


def build_ensemble(nscen=512, seed=1, kmin=32, kmax=700, cell=10.0):
    """Scenario list of the batched workload: k ~ log-uniform[kmin, kmax]
    (numpy default_rng(seed)), rectangular_cross(k, k) over (cell*k)^2 m
    (4k^2 triangles), bed -0.005 x plus 8 Gaussian hills (amplitude
    U(0.5, 3) m, sigma U(0.03, 0.1) L), a hump of U(0.5, 5) m (sigma 0.1 L)
    at a random centre clipped to the bed, Manning n ~ U(0.01, 0.06),
    transmissive west, reflective elsewhere. Returns (k, params) per scenario;
    meshes are built on demand by ensemble_scenario."""
    rng = np.random.default_rng(seed)
    ks = np.rint(np.exp(rng.uniform(np.log(kmin), np.log(kmax), nscen))).astype(int)
    out = []
    for k in ks:
        L = cell * k
        hills = [(rng.uniform(0.5, 3.0), rng.uniform(0.0, L), rng.uniform(0.0, L),
                  rng.uniform(0.03, 0.1) * L) for _ in range(8)]
        hump = (rng.uniform(0.5, 5.0), rng.uniform(0.0, L), rng.uniform(0.0, L))
        out.append((int(k), dict(L=L, hills=hills, hump=hump, n=rng.uniform(0.01, 0.06))))
    return out


def ensemble_scenario(k, prm, mesh=None):
    """Mesh, centroid quantities and boundary kinds of one batched scenario."""
    from anuga_core_amd.mesh import Mesh, rectangular_cross
    L = prm["L"]
    if mesh is None:
        mesh = Mesh(*rectangular_cross(k, k, L, L))
    x = mesh.centroid_coordinates[:, 0]
    y = mesh.centroid_coordinates[:, 1]
    z = -0.005 * x
    for a, cx, cy, s in prm["hills"]:
        z = z + a * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2.0 * s * s))
    H, hx, hy = prm["hump"]
    s = 0.1 * L
    w = np.maximum(H * np.exp(-((x - hx) ** 2 + (y - hy) ** 2) / (2.0 * s * s)), z)
    q = dict(elevation=z, stage=w, friction=np.full(x.shape, prm["n"]))
    bc = {"left": "transmissive", "right": "reflective", "top": "reflective",
          "bottom": "reflective"}
    return mesh, q, bc


def run_ensemble(args):
    """The batched workload: every scenario advanced --steps DE1 steps with its
    own timestep, all scenarios in one device handle."""
    from anuga_core_amd.ensemble import Ensemble
    t0 = time.perf_counter()
    spec = build_ensemble(args.scenarios)
    order = sorted(range(len(spec)), key=lambda i: spec[i][0])
    sizes = [(4 * k * k, 4 * k) for k, _ in spec]
    ens = Ensemble(sizes, flow_algorithm="DE1", precision=args.precision,
                   reorder=_reorder(args))
    mesh, mesh_k = None, None
    for j, i in enumerate(order):
        k, prm = spec[i]
        if k != mesh_k:
            mesh, mesh_k = None, k
        mesh, q, bc = ensemble_scenario(k, prm, mesh)
        ens.set_scenario(i, mesh, q, bc)
        if j % 64 == 63:
            print("  %d scenarios set up (k <= %d) %.1fs" % (j + 1, k, time.perf_counter() - t0),
                  file=sys.stderr, flush=True)
    del mesh
    Ntot = int(sum(n for n, _ in sizes))
    print("ensemble of %d scenarios, %d triangles, set up in %.1fs" %
          (len(spec), Ntot, time.perf_counter() - t0), file=sys.stderr, flush=True)
    v0 = ens.get_water_volumes()
    ens.evolve_steps(args.warmup)
    ens.synchronize()
    t1 = time.perf_counter()
    ens.evolve_steps(args.steps)
    ens.synchronize()
    elapsed = time.perf_counter() - t1
    ens.reset_kernel_timing(True)
    ens.evolve_steps(args.timing_steps)
    ens.synchronize()
    kt = {n: ens.kernel_timing(n) for n in KERNEL_BYTES[args.precision]}
    ens.reset_kernel_timing(False)
    v1 = ens.get_water_volumes()
    bfi = ens.get_boundary_flux_integrals()
    drift = np.abs(v1 - v0 - bfi) / v0
    kernels = {}
    for short, (ms, n) in kt.items():
        if n == 0:
            continue
        kname, nbytes = KERNEL_BYTES[args.precision][short]
        kernels[kernel_label(kname)] = {
            "avg_ms": round(ms, 5), "launches": n, "alg_bytes_per_triangle": nbytes,
            "alg_GBs": round(nbytes * Ntot / (ms * 1e-3) / 1e9, 1) if ms > 0 else None}
    dom_short = max((s for s in kt if kt[s][1] > 0), key=lambda s: kt[s][0] * LAUNCHES_PER_STEP[s])
    dom = kernel_label(KERNEL_BYTES[args.precision][dom_short][0])
    dom_bytes = KERNEL_BYTES[args.precision][dom_short][1]
    ach = kernels[dom]["alg_GBs"]
    ms_per_step = 1e3 * elapsed / args.steps
    value = Ntot * args.steps / elapsed / 1e6
    step_gbs = STEP_BYTES[args.precision] * Ntot / (ms_per_step * 1e-3) / 1e9
    times = ens.get_times()
    out = {
        "metric": "M triangle-steps/s, batched DE1 ensemble (per-scenario dt)",
        "value": round(value, 3), "unit": "M triangle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-flux/f64-state",
        "data": "synthetic",
        "config": {"workload": "%d scenarios rectangular_cross(k,k), k log-uniform[32,700] "
                               "(rng seed 1), 10 m cells, DE1, bed -0.005x + 8 hills, "
                               "hump U(0.5,5) m, n U(0.01,0.06)" % len(spec),
                   "scenarios": len(spec), "triangles": Ntot, "precision": args.precision},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                     "alg_bytes_per_launch": dom_bytes * Ntot},
        "step_roofline": {"alg_bytes_per_triangle_step": STEP_BYTES[args.precision],
                          "achieved": round(step_gbs, 1), "unit": "GB/s",
                          "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
        "kernels": kernels,
        "simulated_time_s": {"min": float(times.min()), "max": float(times.max())},
        "max_volume_drift": float(drift.max()),
    }
    ens.close()
    if not args.no_cpu and args.cpu_steps > 0:
        # bounded sample: the oracle on a few scenarios of the ensemble
        from oracle.kernels import OracleKernels
        from oracle.oracle_domain import OracleDomain
        pick = [order[len(order) // 4], order[len(order) // 2], order[3 * len(order) // 4]]
        # steps so that the sample is about 3e8 triangle-steps (10-30 s on one host)
        nsample = sum(4 * spec[i][0] ** 2 for i in pick)
        csteps = max(args.cpu_steps, int(3.0e8 / nsample))
        tri_steps, secs = 0, 0.0
        for i in pick:
            k, prm = spec[i]
            mesh, q, bc = ensemble_scenario(k, prm)
            o = OracleDomain(mesh, OracleKernels(gather=True), "DE1")
            for name, val in q.items():
                o.set_centroid_values(name, val)
            o.set_boundary({t: type("B", (), {"kind": kk})() for t, kk in bc.items()})
            gen = o.evolve(yieldstep=1.0e300, finaltime=1.0e300, max_steps=csteps)
            next(gen)
            ts = time.perf_counter()
            for _ in gen:
                pass
            secs += time.perf_counter() - ts
            tri_steps += mesh.number_of_triangles * len(o.dt_history)
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = {
            "value": round(tri_steps / secs / 1e6, 4), "unit": "M triangle-steps/s",
            "cores": threads, "kind": "port",
            "sample": "%d DE1 steps of 3 scenarios of the ensemble (k = %s), oracle C "
                      "restatement, %.1f s" % (csteps, [spec[i][0] for i in pick], secs)}
    print(json.dumps(out), flush=True)


def build_c4(m=500, length=5000.0, seed=2, west="transmissive"):
    """The gradient workload: rectangular_cross(m, m) over a 5 km square (10 m
    cells), fully wet (depth >= 5 m): bed -12 m with a 4 m Gaussian bump
    (sigma 800 m), a 0.5 m stage hump (sigma 300 m) off-centre, at rest;
    transmissive west, reflective elsewhere. The true Manning n is 0.03 +
    0.01 * a smooth field in [-1, 1] (sum of 6 random planar waves, numpy
    default_rng(seed)); 32 gauge triangles drawn with the same generator."""
    from anuga_core_amd.mesh import Mesh, rectangular_cross
    mesh = Mesh(*rectangular_cross(m, m, length, length))
    x = mesh.centroid_coordinates[:, 0]
    y = mesh.centroid_coordinates[:, 1]
    z = -12.0 + 4.0 * np.exp(-((x - 0.55 * length) ** 2 + (y - 0.45 * length) ** 2)
                             / (2.0 * 800.0 ** 2))
    w = 0.5 * np.exp(-((x - 0.3 * length) ** 2 + (y - 0.6 * length) ** 2) / (2.0 * 300.0 ** 2))
    rng = np.random.default_rng(seed)
    field = np.zeros_like(x)
    for _ in range(6):
        kx, ky = rng.uniform(-3, 3, 2) * 2 * np.pi / length
        field += np.sin(kx * x + ky * y + rng.uniform(0, 2 * np.pi))
    field /= np.abs(field).max()
    n_true = 0.03 + 0.01 * field
    gauges = rng.choice(mesh.number_of_triangles, 32, replace=False)
    q = dict(elevation=z, stage=w, friction=n_true)
    bc = {"left": west, "right": "reflective", "top": "reflective", "bottom": "reflective"}
    return mesh, q, bc, gauges


def run_adjoint(args):
    """Gradient of a gauge misfit with respect to friction and bed by the
    discrete adjoint, checked against central differences of the device
    forward along random directions.

    The default 2000 steps of 0.025 s cover 50 s. A perturbation of the bed
    of 1e-10 grows by about e every 10 s of simulated time in this wave field
    (limiter switching), so at 200 s (dt 0.1) any perturbation, however
    small, changes the gauge record by about 1e-5 and neither central
    differences nor the derivative of the discrete map are meaningful."""
    from anuga_core_amd.adjoint import GaugeRun
    from anuga_core_amd.domain import Domain
    t0 = time.perf_counter()
    mesh, q, bc, gauges = build_c4(west=args.adjoint_west)
    N = mesh.number_of_triangles
    d = Domain(mesh=mesh, reorder=_reorder(args))
    d.set_flow_algorithm("DE1")
    for name, val in q.items():
        d.set_quantity(name, val, location="centroids")
    d.set_boundary(_boundaries(d, bc))
    init = dict(stage=q["stage"], xmomentum=np.zeros(N), ymomentum=np.zeros(N))
    nsteps, dt, obs_every, every = args.steps, args.adjoint_dt, 10, 100
    run = GaugeRun(d, init, nsteps, dt, gauges, obs_every, every)
    print("gradient workload ready: %d triangles in %.1fs" % (N, time.perf_counter() - t0),
          file=sys.stderr, flush=True)
    obs = run.forward(q["friction"], q["elevation"])          # synthetic observations
    n0 = np.full(N, 0.03)
    z0 = q["elevation"]
    d.synchronize()
    t1 = time.perf_counter()
    rec = run.forward(n0, z0)
    d.synchronize()
    t_fwd = time.perf_counter() - t1
    loss0 = float(np.sum((rec - obs) ** 2))
    t2 = time.perf_counter()
    gn, gz, gs = d.adjoint_gradient(2.0 * (rec - obs))
    t_adj = time.perf_counter() - t2
    print("forward %.2fs, gradient %.2fs, loss %.6e" % (t_fwd, t_adj, loss0), file=sys.stderr,
          flush=True)

    def loss(n, z):
        return float(np.sum((run.forward(n, z) - obs) ** 2))

    rng = np.random.default_rng(7)
    fd_eps = [float(e) for e in args.fd_eps.split(",")]
    checks = []
    for j in range(args.fd_directions):
        which = "friction" if j % 2 == 0 else "elevation"
        dvec = rng.standard_normal(N)
        dvec /= np.linalg.norm(dvec)
        ad = float(np.dot(gn if which == "friction" else gz, dvec))
        row = {"parameter": which, "adjoint": ad}
        for eps in fd_eps:
            if which == "friction":
                fd = (loss(n0 + eps * dvec, z0) - loss(n0 - eps * dvec, z0)) / (2 * eps)
            else:
                fd = (loss(n0, z0 + eps * dvec) - loss(n0, z0 - eps * dvec)) / (2 * eps)
            row["fd_%g" % eps] = fd
            row["rel_%g" % eps] = abs(ad - fd) / max(abs(fd), 1e-300)
        checks.append(row)
        print("direction %d %s adjoint %.9e fd %s" %
              (j, which, ad, " ".join("(%g) %.9e" % (e, row["fd_%g" % e]) for e in fd_eps)),
              file=sys.stderr, flush=True)
    best = [min(r["rel_%g" % e] for e in fd_eps) for r in checks]
    out = {
        "metric": "seconds for the discrete-adjoint gradient of a 2000-step DE1 gauge loss "
                  "(friction and bed per triangle)",
        "value": round(t_adj, 3), "unit": "s", "n_gpus": 1, "steps": nsteps, "warmup": 0,
        "ms_per_step": round(1e3 * t_adj / nsteps, 3), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "rectangular_cross(500,500) 5 km, DE1 fixed dt %.2f s, fully wet, "
                               "32 gauges every %d steps, true n = 0.03 + 0.01 smooth field "
                               "(seed 2), start n = 0.03, checkpoint every %d steps, %s west, "
                               "reflective elsewhere"
                               % (dt, obs_every, every, args.adjoint_west),
                   "triangles": N},
        "forward_seconds": round(t_fwd, 3),
        "gradient_over_forward": round(t_adj / t_fwd, 2),
        "loss": loss0,
        "grad_norm": {"friction": float(np.linalg.norm(gn)), "elevation": float(np.linalg.norm(gz))},
        "fd_check": {"directions": len(checks), "eps": fd_eps,
                     "median_rel": float(np.median(best)) if best else None,
                     "max_rel": float(np.max(best)) if best else None,
                     "rows": checks},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: 1000 single, 500 ensemble, 2000 adjoint)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--m", type=int, default=2000)
    ap.add_argument("--mesh", default="rectangular_cross", choices=["rectangular_cross", "graded"],
                    help="single workload mesh: rectangular_cross(m, m), or the unstructured "
                         "graded variant (about 4M triangles)")
    ap.add_argument("--cpu-steps", type=int, default=100,
                    help="steps of the CPU baseline on all the host's OpenMP threads")
    ap.add_argument("--cpu-steps-1t", type=int, default=4,
                    help="steps of the CPU baseline on one thread")
    ap.add_argument("--timing-steps", type=int, default=50,
                    help="steps of the eager pass that times each kernel with HIP events")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the second precision (speed and mixed-vs-FP64 error)")
    ap.add_argument("--order", default="tiles", choices=["tiles", "rcm", "none"],
                    help="internal triangle order: compact tiles (fused FP64 kernel), "
                         "Reverse Cuthill-McKee, or the mesh's own")
    ap.add_argument("--no-reorder", action="store_true", help="same as --order none")
    ap.add_argument("--workload", default="single", choices=["single", "ensemble", "adjoint"],
                    help="single: the 16M-triangle mesh; ensemble: the batched scenarios; "
                         "adjoint: gradients on the 1M-triangle calibration case")
    ap.add_argument("--fd-directions", type=int, default=20)
    ap.add_argument("--adjoint-west", default="transmissive",
                    choices=["transmissive", "reflective"],
                    help="west boundary of the gradient workload (the others are reflective)")
    ap.add_argument("--adjoint-dt", type=float, default=0.025,
                    help="fixed timestep (s) of the gradient workload; beyond about 60 s "
                         "simulated the run is chaotic at rounding level and central "
                         "differences no longer converge")
    ap.add_argument("--fd-eps", default="1e-6,1e-7",
                    help="central-difference steps of the gradient check (comma list)")
    ap.add_argument("--scenarios", type=int, default=512)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run 3 device copies of 1 GiB (a known-traffic kernel for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = {"single": 1000, "ensemble": 500, "adjoint": 2000}[args.workload]
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the step shards by spatial domain decomposition, not implemented here)",
              file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        import torch
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    if args.workload == "ensemble":
        run_ensemble(args)
        return
    if args.workload == "adjoint":
        run_adjoint(args)
        return

    t0 = time.perf_counter()
    mesh, q, bc = build_c1(args.m, kind=args.mesh)
    N = mesh.number_of_triangles
    print("mesh built: %d triangles in %.1fs; dry fraction %.3f" %
          (N, time.perf_counter() - t0, float(np.mean(q["stage"] - q["elevation"] <= 1e-5))),
          file=sys.stderr, flush=True)

    d, r = run_timed(args.precision, mesh, q, bc, args)
    kernels = r["kernels"]
    # dominant kernel: largest total time per step
    label = {s: kernel_label(v[0]) for s, v in KERNEL_BYTES[args.precision].items()}
    ran = [s for s in KERNEL_BYTES[args.precision] if label[s] in kernels]
    dom_short = max(ran, key=lambda s: kernels[label[s]]["avg_ms"] * LAUNCHES_PER_STEP[s])
    dom, dom_bytes = label[dom_short], KERNEL_BYTES[args.precision][dom_short][1]
    ach = kernels[dom]["alg_GBs"]
    traffic = measured_traffic(args.precision, dom, args.m, args.mesh)
    step_gbs = STEP_BYTES[args.precision] * N / (r["ms_per_step"] * 1e-3) / 1e9

    out = {
        "metric": "M triangle-steps/s at 16M-tri DE1: FP64 vs FP32-flux (stated err); HBM GB/s",
        "value": round(r["value"], 3),
        "unit": "M triangle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-flux/f64-state",
        "data": "synthetic",
        "config": {"workload": ("rectangular_cross(%d,%d)" % (args.m, args.m) if args.mesh != "graded"
                                else "graded Delaunay mesh (50 m^2 in the central 13 km square, "
                                     "400 m^2 outside, seed 0)")
                               + " over %.0f km square, DE1, 64 Gaussian "
                               "hills (rng seed 0), 3 m hump, n=0.025, transmissive west"
                               % 20.0,
                   "triangles": N, "flow_algorithm": "DE1", "precision": args.precision,
                   "order": args.order},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": ach,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "alg_bytes_per_launch": dom_bytes * N},
        "step_roofline": {"alg_bytes_per_triangle_step": STEP_BYTES[args.precision],
                          "achieved": round(step_gbs, 1), "unit": "GB/s",
                          "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
        "kernels": kernels,
        "volume_drift": r["volume_drift"],
    }

    if not args.no_compare:
        other = "mixed" if args.precision == "fp64" else "fp64"
        d2, r2 = run_timed(other, mesh, q, bc, args)
        mixed, ref = (d2, d) if other == "mixed" else (d, d2)
        out[other] = {"value": round(r2["value"], 3), "ms_per_step": round(r2["ms_per_step"], 4),
                      "kernels": r2["kernels"], "volume_drift": r2["volume_drift"]}
        rv = {"fp64": r if args.precision == "fp64" else r2,
              "mixed": r2 if args.precision == "fp64" else r}
        out["mixed_speedup_vs_fp64"] = round(rv["mixed"]["value"] / rv["fp64"]["value"], 3)
        out["mixed_error_vs_fp64"] = dict(error_report(mixed, ref),
                                          steps=r["steps_total"])
        d2.close()

    if not args.no_cpu and args.cpu_steps > 0:
        nsteps, dt, threads = cpu_baseline(mesh, q, bc, args.cpu_steps)
        out["cpu_baseline"] = {
            "value": round(N * nsteps / dt / 1e6, 4), "unit": "M triangle-steps/s",
            "cores": threads, "kind": "port",
            "sample": "%d DE1 steps of the same %d-triangle workload, oracle C restatement "
                      "(OpenMP per-triangle loops, %d threads), %.1f s"
                      % (nsteps, N, threads, dt),
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count()}
        if args.cpu_steps_1t > 0:
            n1, dt1, _ = cpu_baseline(mesh, q, bc, args.cpu_steps_1t, threads=1)
            out["cpu_baseline"]["one_thread"] = {
                "value": round(N * n1 / dt1 / 1e6, 4), "unit": "M triangle-steps/s",
                "cores": 1, "sample": "%d DE1 steps of the same workload, 1 thread, %.1f s"
                                      % (n1, dt1)}
        out["speedup_vs_cpu"] = round(r["value"] / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Benchmark: M triangle-steps/s of the DE1 step on a 16M-triangle mesh.

    python bench.py [--gpus 1] [--steps 1000] [--warmup 20] [--precision fp64|mixed]
                    [--m 2000] [--cpu-steps 3] [--no-cpu]

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

# Algorithmic bytes per triangle (FP64), each input of the unit read once and
# each output written once:
#   whole DE1 step: 2 x 232 B per Euler substep + 24 B for reading the
#   step's initial state in the final saxpy = 488 B;
#   per kernel (this implementation's split of the substep):
KERNEL_BYTES_FP64 = {
    # w,uh,vh,z (32) + neighbours (12) + centroid (16) + edge offsets (48)
    #   -> 4 edge quantities x 3 edges (96) + flag (1)
    "de_extrapolate": 205,
    # edge values (96) + w,z (16) + flag (1) + neighbours (12) + normals (48)
    #   + lengths (24) + area, radius (16) -> explicit updates (24)
    "de_flux<0>": 237,
    # w,uh,vh,z,n (40) + flag (1) + explicit updates (24) -> w,uh,vh (24)
    "de_update": 89,
    # de_flux<0> inputs + uh,vh,n (24) + backup state (24) -> state (24)
    "de_flux<1>": 277,
}
STEP_BYTES = {"fp64": 488, "mixed": 344}


def build_c1(m=2000, length=20000.0, nhills=64, seed=0):
    """Mesh and initial condition of the 16M-triangle DE1 workload."""
    from anuga_core_amd.mesh import Mesh, rectangular_cross
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


def cpu_baseline(mesh, q, bc, steps):
    """The oracle (C restatement of the reference's FP64 path, OpenMP per-
    triangle flux loop, libm pow as the reference) on this host, timed over a
    bounded number of steps of the same workload."""
    from oracle.kernels import OracleKernels
    from oracle.oracle_domain import OracleDomain
    o = OracleDomain(mesh, OracleKernels(gather=True, pow_mode=0), "DE1")
    for name, val in q.items():
        o.set_centroid_values(name, val)
    o.set_boundary({t: type("B", (), {"kind": k})() for t, k in bc.items()})
    gen = o.evolve(yieldstep=1.0e300, finaltime=1.0e300, max_steps=steps)
    next(gen)   # the initial yield comes before any step
    t0 = time.perf_counter()
    for _ in gen:
        pass
    dt = time.perf_counter() - t0
    nsteps = len(o.dt_history)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return nsteps, dt, threads


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--m", type=int, default=2000)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--timing-steps", type=int, default=50,
                    help="steps of the eager pass that times each kernel with HIP events")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-reorder", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run 3 device copies of 1 GiB (a known-traffic kernel for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the step shards by spatial domain decomposition, not implemented here)",
              file=sys.stderr)
        sys.exit(2)

    from anuga_core_amd.domain import Domain

    if args.calibrate_copy:
        import torch
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    t0 = time.perf_counter()
    mesh, q, bc = build_c1(args.m)
    N = mesh.number_of_triangles
    print("mesh built: %d triangles in %.1fs" % (N, time.perf_counter() - t0),
          file=sys.stderr, flush=True)
    d = Domain(mesh=mesh, precision=args.precision, reorder=not args.no_reorder)
    d.set_flow_algorithm("DE1")
    for name, val in q.items():
        d.set_quantity(name, val, location="centroids")
    d.set_boundary(_boundaries(d, bc))
    d.record_dt_history = False
    print("device domain ready %.1fs; dry fraction %.3f" %
          (time.perf_counter() - t0, float(np.mean(q["stage"] - q["elevation"] <= 1e-5))),
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
    # per-kernel durations: HIP events around every launch on the library's
    # stream, over an eager (non-graph) run of the same steps
    d.reset_kernel_timing(True)
    d.evolve_steps(args.timing_steps)
    d.synchronize()
    kt = {n: d.kernel_timing(n) for n in ("extrapolate", "flux0", "update", "flux1")}
    d.reset_kernel_timing(False)
    d.distribute_to_vertices_and_edges()
    v1 = d.get_water_volume()
    bfi = d.boundary_flux_integral

    ms_per_step = 1e3 * elapsed / args.steps
    value = N * args.steps / elapsed / 1e6
    names = {"extrapolate": "de_extrapolate", "flux0": "de_flux<0>", "update": "de_update",
             "flux1": "de_flux<1>"}
    per_kernel = {}
    for short, (ms, n) in kt.items():
        kname = names[short]
        b = KERNEL_BYTES_FP64[kname] * N
        per_kernel[kname] = {"avg_ms": ms, "launches": n,
                             "alg_GBs": (b / (ms * 1e-3) / 1e9) if ms > 0 else None}
    dom = max(per_kernel, key=lambda k: per_kernel[k]["avg_ms"] * per_kernel[k]["launches"])
    ach = per_kernel[dom]["alg_GBs"]
    step_bytes = STEP_BYTES[args.precision] * N
    step_gbs = step_bytes / (ms_per_step * 1e-3) / 1e9

    out = {
        "metric": "M triangle-steps/s at 16M-tri DE1: FP64 vs FP32-flux (stated err); HBM GB/s",
        "value": round(value, 3),
        "unit": "M triangle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-flux/f64-state",
        "data": "synthetic",
        "config": {"workload": "rectangular_cross(%d,%d) 20km DE1, 64 Gaussian hills seed 0, "
                               "3 m hump, n=0.025, transmissive west" % (args.m, args.m),
                   "triangles": N, "flow_algorithm": "DE1", "precision": args.precision,
                   "reorder": "rcm" if not args.no_reorder else "none"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                     "alg_bytes_per_triangle": KERNEL_BYTES_FP64[dom]},
        "step_roofline": {"alg_bytes_per_triangle_step": STEP_BYTES[args.precision],
                          "achieved": round(step_gbs, 1), "unit": "GB/s",
                          "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
        "kernels": {k: {kk: (round(vv, 5) if isinstance(vv, float) else vv)
                        for kk, vv in v.items()} for k, v in per_kernel.items()},
        "volume_drift": (v1 - v0 - bfi) / v0,
    }

    if not args.no_cpu and args.cpu_steps > 0:
        nsteps, dt, threads = cpu_baseline(mesh, q, bc, args.cpu_steps)
        out["cpu_baseline"] = {
            "value": round(N * nsteps / dt / 1e6, 4), "unit": "M triangle-steps/s",
            "cores": threads, "kind": "port",
            "sample": "%d DE1 steps of the same 16M-triangle workload, oracle C restatement "
                      "(OpenMP flux loop), %.1f s" % (nsteps, dt)}
        out["speedup_vs_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Full-size C1 parity against the oracle (a script, not collected by pytest).

The 16M-triangle DE1 workload of bench.py (rectangular_cross(2000,2000) over
20 km, 64 hills, 3 m hump, transmissive west), STEPS steps on the device
(default internal order: the fused FP64 kernels) and on the oracle (the C
library's pow, as the reference's C calls it; the device's pow returns the
same bits).

The two runs can be made on different machines, each writing a summary:

    python tests/validate_c1_full.py --part device --out dev.npz   # the GPU box
    python tests/validate_c1_full.py --part oracle --out ora.npz   # any host
    python tests/validate_c1_full.py --compare dev.npz ora.npz

A summary holds the timestep sequence, the volume balance, a SHA-256 digest
of each final centroid array and of each of its 4096 contiguous chunks, and
every 97th triangle's values (for a max-norm estimate where bits differ).
`--part both` runs both here and compares the full arrays.
"""

import argparse
import hashlib
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

NAMES = ("stage", "xmomentum", "ymomentum")
CHUNKS = 4096
STRIDE = 97


def _summary(arrays, dt, drift, seconds):
    out = {"dt": np.asarray(dt, np.float64), "drift": np.float64(drift),
           "seconds": np.float64(seconds)}
    for name in NAMES:
        a = np.ascontiguousarray(arrays[name], np.float64)
        out[name + "_sha"] = np.frombuffer(hashlib.sha256(a.tobytes()).digest(), np.uint8)
        out[name + "_chunks"] = np.array(
            [np.frombuffer(hashlib.sha256(c.tobytes()).digest()[:8], np.uint64)[0]
             for c in np.array_split(a, CHUNKS)], np.uint64)
        out[name + "_sample"] = a[::STRIDE].copy()
    return out


def run_device(mesh, q, bc, steps):
    import bench
    from anuga_core_amd.domain import Domain
    t0 = time.perf_counter()
    d = Domain(mesh=mesh)
    d.set_flow_algorithm("DE1")
    for name, val in q.items():
        d.set_quantity(name, val, location="centroids")
    d.set_boundary(bench._boundaries(d, bc))
    d.distribute_to_vertices_and_edges()
    v0 = d.get_water_volume()
    d.evolve_steps(steps)
    d.distribute_to_vertices_and_edges()
    arrays = {k: d.get_quantity(k).centroid_values.copy() for k in NAMES}
    dt = list(d.dt_history)
    drift = (d.get_water_volume() - v0 - d.boundary_flux_integral) / v0
    d.close()
    print("device: %d steps, t = %.6f s, volume drift %.2e" % (len(dt), sum(dt), drift),
          flush=True)
    return arrays, dt, drift, time.perf_counter() - t0


def run_oracle(mesh, q, bc, steps):
    from oracle.kernels import OracleKernels
    from oracle.oracle_domain import OracleDomain
    o = OracleDomain(mesh, OracleKernels(gather=True), "DE1")
    for name, val in q.items():
        o.set_centroid_values(name, val)
    o.set_boundary({t: type("B", (), {"kind": k})() for t, k in bc.items()})
    step = o.evolve_one_rk2_step
    count = [0]
    t1 = time.perf_counter()

    def counted(*args):
        step(*args)
        count[0] += 1
        if count[0] % 50 == 0:
            print("oracle step %d (%.0fs)" % (count[0], time.perf_counter() - t1), flush=True)
    o.evolve_one_rk2_step = counted
    o.distribute_to_vertices_and_edges()
    v0 = o.get_water_volume()
    for _ in o.evolve(yieldstep=1.0e300, finaltime=1.0e300, max_steps=steps):
        pass
    o.distribute_to_vertices_and_edges()
    drift = (o.get_water_volume() - v0 - o.boundary_flux_integral) / v0
    arrays = {"stage": o.stage_c.copy(), "xmomentum": o.xmom_c.copy(),
              "ymomentum": o.ymom_c.copy()}
    seconds = time.perf_counter() - t1
    print("oracle: %d steps in %.0fs, volume drift %.2e" % (len(o.dt_history), seconds, drift),
          flush=True)
    return arrays, list(o.dt_history), drift, seconds


def compare(a, b):
    """Print the comparison of two summaries; returns True when bitwise."""
    same_dt = np.array_equal(a["dt"], b["dt"])
    print("timesteps: %d and %d, sequences identical: %s" % (len(a["dt"]), len(b["dt"]), same_dt))
    if not same_dt and len(a["dt"]) and len(b["dt"]):
        n = min(len(a["dt"]), len(b["dt"]))
        diff = np.nonzero(a["dt"][:n] != b["dt"][:n])[0]
        if len(diff):
            print("  first differing step %d" % diff[0])
    print("volume drift: %.3e and %.3e" % (a["drift"], b["drift"]))
    bitwise = same_dt
    for name in NAMES:
        same = np.array_equal(a[name + "_sha"], b[name + "_sha"])
        chunks = float(np.mean(a[name + "_chunks"] == b[name + "_chunks"]))
        x, y = a[name + "_sample"], b[name + "_sample"]
        rel = np.max(np.abs(x - y)) / max(np.max(np.abs(y)), 1e-300)
        print("%-9s bitwise %s, identical chunks %.4f, sampled rel max-norm %.3e"
              % (name, same, chunks, rel))
        bitwise = bitwise and same
    print("BITWISE" if bitwise else "NOT BITWISE", flush=True)
    return bitwise


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--m", type=int, default=2000)
    ap.add_argument("--part", choices=["device", "oracle", "both"], default=None)
    ap.add_argument("--out", default=None)
    ap.add_argument("--compare", nargs=2, default=None)
    a = ap.parse_args()
    if a.compare:
        x, y = (dict(np.load(f)) for f in a.compare)
        sys.exit(0 if compare(x, y) else 1)
    import bench
    t0 = time.perf_counter()
    mesh, q, bc = bench.build_c1(a.m)
    print("workload: %d triangles (%.0fs)" % (mesh.number_of_triangles, time.perf_counter() - t0),
          flush=True)
    part = a.part or "both"
    res = {}
    if part in ("device", "both"):
        res["device"] = run_device(mesh, q, bc, a.steps)
    if part in ("oracle", "both"):
        res["oracle"] = run_oracle(mesh, q, bc, a.steps)
    if part == "both":
        (x, dx, _, _), (y, dy, _, _) = res["device"], res["oracle"]
        print("timestep sequences identical: %s" % (dx == dy), flush=True)
        for name in NAMES:
            rel = np.max(np.abs(x[name] - y[name])) / max(np.max(np.abs(y[name])), 1e-300)
            print("%-9s rel max-norm %.3e, bit-identical triangles %.6f"
                  % (name, rel, float(np.mean(x[name] == y[name]))), flush=True)
    elif a.out:
        np.savez(a.out, **_summary(*res[part]))
        print("summary written to %s" % a.out, flush=True)


if __name__ == "__main__":
    main()

"""Full-size C1 parity against the oracle (a script, not collected by pytest:
about 12 minutes, most of it the oracle on the host's threads).

The 16M-triangle DE1 workload of bench.py (rectangular_cross(2000,2000) over
20 km, 64 hills, 3 m hump, transmissive west), STEPS steps on the device
(default internal order: the fused FP64 kernels) and on the oracle (the C library's pow, as
the reference's C calls it; the device's pow returns the same bits). Prints whether the
timestep sequences agree, the relative max-norm of stage and momenta, the
share of bit-identical triangles and both volume balances.

    python tests/validate_c1_full.py [--steps 1000] [--m 2000]
"""

import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--m", type=int, default=2000)
    a = ap.parse_args()
    import bench
    from anuga_core_amd.domain import Domain
    from oracle.kernels import OracleKernels
    from oracle.oracle_domain import OracleDomain

    t0 = time.perf_counter()
    mesh, q, bc = bench.build_c1(a.m)
    N = mesh.number_of_triangles
    print("workload: %d triangles (%.0fs)" % (N, time.perf_counter() - t0), flush=True)

    d = Domain(mesh=mesh)
    d.set_flow_algorithm("DE1")
    for name, val in q.items():
        d.set_quantity(name, val, location="centroids")
    d.set_boundary(bench._boundaries(d, bc))
    d.distribute_to_vertices_and_edges()
    v0 = d.get_water_volume()
    d.evolve_steps(a.steps)
    d.distribute_to_vertices_and_edges()
    dev = {k: d.get_quantity(k).centroid_values.copy() for k in ("stage", "xmomentum", "ymomentum")}
    dev_dt = list(d.dt_history)
    dev_drift = (d.get_water_volume() - v0 - d.boundary_flux_integral) / v0
    d.close()
    print("device: %d steps, t = %.6f s, volume drift %.2e" % (len(dev_dt), sum(dev_dt), dev_drift),
          flush=True)

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
    ov0 = o.get_water_volume()
    for _ in o.evolve(yieldstep=1.0e300, finaltime=1.0e300, max_steps=a.steps):
        pass
    o.distribute_to_vertices_and_edges()
    o_drift = (o.get_water_volume() - ov0 - o.boundary_flux_integral) / ov0
    ora = {"stage": o.stage_c, "xmomentum": o.xmom_c, "ymomentum": o.ymom_c}
    print("oracle: %d steps in %.0fs, volume drift %.2e" % (len(o.dt_history),
                                                          time.perf_counter() - t1, o_drift),
          flush=True)

    print("timestep sequences identical: %s" % (dev_dt == list(o.dt_history)), flush=True)
    for name in ("stage", "xmomentum", "ymomentum"):
        x, y = dev[name], ora[name]
        rel = np.max(np.abs(x - y)) / max(np.max(np.abs(y)), 1e-300)
        print("%-9s rel max-norm %.3e, bit-identical triangles %.6f" %
              (name, rel, float(np.mean(x == y))), flush=True)


if __name__ == "__main__":
    main()

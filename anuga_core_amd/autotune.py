"""Autotuner of the library's compile-time block (tile) size and of the mixed
fused kernel's occupancy target.

    python -m anuga_core_amd.autotune build     # compile every variant (hipcc, no GPU)
    python -m anuga_core_amd.autotune measure   # time them on the GPU, write tuned.json

Variants: DE_BLOCK in {128, 256, 512} (threads per workgroup, which is also the
number of triangles per tile of the fused extrapolation + flux kernels) times
DE_MIXED_WAVES in {4, 5, 6} (the waves per SIMD the mixed fused kernel's register
allocation aims at). `measure` runs each variant in a child process on a
DE1 workload of rectangular_cross(m, m) at 10 m cells (bed -50 m plus 8
hills, a 3 m hump, n = 0.025, transmissive west; numpy default_rng(0)), FP64
and mixed, and chooses the block size by the FP64 step time and then the
wave target by the mixed step time at that block size. tuned.json records the
measurements and the choice; build() compiles with it. Results are the same
bits for every variant (only per-workgroup partial sums, such as the boundary
flux integral, round differently).
"""

from __future__ import annotations

import argparse
import itertools
import json
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
VARIANTS_DIR = os.path.join(HERE, "_variants")
BLOCKS = (128, 256, 512)
WAVES = (4, 5, 6)


def _variant_path(block, waves):
    return os.path.join(VARIANTS_DIR, "libanuga_hip_b%d_w%d.so" % (block, waves))


def build_variants(verbose=True):
    from anuga_core_amd.build import build
    os.makedirs(VARIANTS_DIR, exist_ok=True)
    for block, waves in itertools.product(BLOCKS, WAVES):
        t0 = time.perf_counter()
        build(out=_variant_path(block, waves),
              config={"DE_BLOCK": block, "DE_MIXED_WAVES": waves})
        if verbose:
            print("built block %d waves %d in %.1fs" % (block, waves, time.perf_counter() - t0),
                  flush=True)


def _workload(m):
    import numpy as np
    from anuga_core_amd.mesh import Mesh, rectangular_cross
    L = 10.0 * m
    mesh = Mesh(*rectangular_cross(m, m, L, L))
    x, y = mesh.centroid_coordinates[:, 0], mesh.centroid_coordinates[:, 1]
    rng = np.random.default_rng(0)
    z = np.full(x.shape, -50.0)
    for _ in range(8):
        a, s = rng.uniform(5.0, 60.0), rng.uniform(0.01, 0.075) * L
        cx, cy = rng.uniform(0, L, 2)
        z = z + a * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * s * s))
    sig = 0.05 * L
    w = np.maximum(3.0 * np.exp(-((x - L / 2) ** 2 + (y - L / 2) ** 2) / (2 * sig * sig)), z)
    return mesh, dict(elevation=z, stage=w, friction=np.full(x.shape, 0.025))


def time_one(precision, m, steps, warmup=20):
    """Milliseconds per DE1 step of the library this process loads."""
    from anuga_core_amd import Reflective_boundary, Transmissive_boundary
    from anuga_core_amd.domain import Domain
    mesh, q = _workload(m)
    d = Domain(mesh=mesh, precision=precision)
    d.set_flow_algorithm("DE1")
    for name, val in q.items():
        d.set_quantity(name, val, location="centroids")
    d.set_boundary({"left": Transmissive_boundary(d), "right": Reflective_boundary(d),
                    "top": Reflective_boundary(d), "bottom": Reflective_boundary(d)})
    d.evolve_steps(warmup)
    d.synchronize()
    t0 = time.perf_counter()
    d.evolve_steps(steps)
    d.synchronize()
    return 1e3 * (time.perf_counter() - t0) / steps, mesh.number_of_triangles


def choose(rows):
    """The block size with the fastest FP64 step, then the wave target with
    the fastest mixed step at that block size."""
    fp64 = [r for r in rows if r["precision"] == "fp64"]
    block = min(fp64, key=lambda r: r["ms_per_step"])["DE_BLOCK"]
    mixed = [r for r in rows if r["precision"] == "mixed" and r["DE_BLOCK"] == block]
    waves = min(mixed, key=lambda r: r["ms_per_step"])["DE_MIXED_WAVES"]
    return {"DE_BLOCK": block, "DE_MIXED_WAVES": waves}


def measure(m=1000, steps=200, timeout=300, out_path=None):
    rows = []
    for block, waves in itertools.product(BLOCKS, WAVES):
        lib = _variant_path(block, waves)
        if not os.path.exists(lib):
            raise RuntimeError("variant %s is not built: run `autotune build` first" % lib)
        for precision in ("fp64", "mixed"):
            if precision == "fp64" and waves != 5:
                continue        # the wave target only changes the mixed kernel
            env = dict(os.environ, ANUGA_HIP_LIBRARY=lib)
            r = subprocess.run([sys.executable, "-m", "anuga_core_amd.autotune", "time",
                                "--precision", precision, "--m", str(m), "--steps", str(steps)],
                               env=env, capture_output=True, text=True, timeout=timeout)
            if r.returncode != 0:
                sys.stderr.write(r.stdout[-2000:] + r.stderr[-2000:])
                raise RuntimeError("timing block %d waves %d %s failed" % (block, waves, precision))
            res = json.loads(r.stdout.strip().splitlines()[-1])
            row = {"DE_BLOCK": block, "DE_MIXED_WAVES": waves, "precision": precision,
                   "ms_per_step": res["ms_per_step"], "triangles": res["triangles"]}
            rows.append(row)
            print(json.dumps(row), flush=True)
    out = {"config": choose(rows),
           "workload": "rectangular_cross(%d,%d) DE1 at 10 m cells, %d timed steps" % (m, m, steps),
           "measured": rows}
    with open(out_path or os.path.join(HERE, "tuned.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print("chosen", json.dumps(out["config"]), flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("stage", choices=["build", "measure", "time"])
    ap.add_argument("--m", type=int, default=1000)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--out", default=None, help="where measure writes tuned.json")
    a = ap.parse_args()
    if a.stage == "build":
        build_variants()
    elif a.stage == "measure":
        measure(a.m, a.steps, out_path=a.out)
    else:
        ms, n = time_one(a.precision, a.m, a.steps)
        print(json.dumps({"ms_per_step": ms, "triangles": n}), flush=True)


if __name__ == "__main__":
    main()

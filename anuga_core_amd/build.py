"""Build the device library in-tree: anuga_core_amd/libanuga_hip.so (gfx950).

    python -m anuga_core_amd.build

FP64 kernels are compiled with -ffp-contract=off so that no multiply-add is
fused, which keeps them bit-comparable with the reference's FP64 C. Packed FP32
instructions are disabled (-packed-fp32-ops).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
OUT = os.path.join(HERE, "libanuga_hip.so")
SOURCES = [os.path.join(HERE, "csrc", "de_host.hip")]
DEPENDS = [os.path.join(HERE, "csrc", f) for f in ("de_host.hip", "de_kernels.hip",
                                                    "de_device.h")] + \
          [os.path.join(ROOT, "include", "anuga_hip.h")]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


TUNED = os.path.join(HERE, "tuned.json")
# compile-time parameters the autotuner chooses (anuga_core_amd/autotune.py)
TUNABLE = {"DE_BLOCK": 256, "DE_MIXED_WAVES": 5}


def tuned_config():
    """The autotuner's recorded choice (tuned.json), else the defaults."""
    import json
    cfg = dict(TUNABLE)
    if os.path.exists(TUNED):
        with open(TUNED) as f:
            chosen = json.load(f).get("config", {})
        cfg.update({k: int(v) for k, v in chosen.items() if k in TUNABLE})
    return cfg


def build(verbose=False, out=None, config=None):
    """Compile the library (to `out`, default in-tree) with `config` (default:
    the tuned configuration)."""
    out = out or OUT
    cfg = tuned_config() if config is None else dict(TUNABLE, **config)
    tmp = out + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-ffp-contract=off", "-Wno-unused-value", "-Wno-unused-result",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"] + \
          ["-D%s=%d" % kv for kv in sorted(cfg.items())] + \
          ["-I" + os.path.join(ROOT, "include"), "-o", tmp] + SOURCES
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building %s" % out)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))

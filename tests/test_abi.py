"""The C ABI library: loads, exports every declared symbol, host-only helpers
(CPU only; no compute call needs a device here)."""

import ctypes as C
import os
import re

import numpy as np
import pytest

from anuga_core_amd import _native
from anuga_core_amd.mesh import Mesh, rectangular_cross

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    src = open(os.path.join(ROOT, "include", "anuga_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(anuga_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = _native.lib()
    names = _declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    # the ctypes signature table covers the header exactly
    assert set(names) == set(_native.SIGNATURES)


def test_rcm_order_is_a_bandwidth_reducing_permutation():
    m = Mesh(*rectangular_cross(30, 20))
    N = m.number_of_triangles
    nb = np.ascontiguousarray(m.neighbours, np.int64)
    order = np.zeros(N, np.int64)
    rc = _native.lib().anuga_rcm_order(N, nb.ctypes.data_as(C.POINTER(C.c_int64)),
                                       order.ctypes.data_as(C.POINTER(C.c_int64)))
    assert rc == 0
    assert np.array_equal(np.sort(order), np.arange(N))
    inv = np.empty_like(order)
    inv[order] = np.arange(N)
    k = np.repeat(np.arange(N), 3)
    n = nb.ravel()
    sel = n >= 0
    bw_nat = np.max(np.abs(k[sel] - n[sel]))
    bw_rcm = np.max(np.abs(inv[k[sel]] - inv[n[sel]]))
    assert bw_rcm < bw_nat
    # invalid input is rejected, not read out of bounds
    bad = nb.copy()
    bad[0, 0] = N + 5
    assert _native.lib().anuga_rcm_order(N, bad.ctypes.data_as(C.POINTER(C.c_int64)),
                                         order.ctypes.data_as(C.POINTER(C.c_int64))) == -1


def test_create_rejects_bad_arguments_before_touching_the_device():
    L = _native.lib()
    h = C.c_void_p()
    assert L.anuga_create(None, None, C.byref(h)) == _native.ANUGA_E_ARG
    m = Mesh(*rectangular_cross(2, 2))
    keep = [np.ascontiguousarray(a) for a in (m.neighbours, m.neighbour_edges,
                                              m.surrogate_neighbours, m.number_of_boundaries,
                                              m.normals, m.edgelengths, m.radii, m.areas,
                                              m.centroid_coordinates,
                                              m.edge_midpoint_coordinates)]
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int64))
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    desc = _native.AnugaMesh(m.number_of_triangles, m.boundary_length, ip(keep[0]), ip(keep[1]),
                             ip(keep[2]), ip(keep[3]), dp(keep[4]), dp(keep[5]), dp(keep[6]),
                             dp(keep[7]), dp(keep[8]), dp(keep[9]))
    p = _native.AnugaParams(9.8, 1e-12, 1e-5, 1, 0, 1, 0, 1, 0, 3, 1, 2, 0, 1)  # low_froude 3
    assert L.anuga_create(C.byref(desc), C.byref(p), C.byref(h)) == _native.ANUGA_E_ARG
    assert b"low_froude" in L.anuga_last_error(None)
    p.low_froude = 1
    p.precision = _native.PRECISION_MIXED     # low_froude needs the FP64 path
    assert L.anuga_create(C.byref(desc), C.byref(p), C.byref(h)) == _native.ANUGA_E_ARG
    assert b"FP64" in L.anuga_last_error(None)
    p.precision = _native.PRECISION_FP64
    p.low_froude = 0
    keep[3][0] += 1   # inconsistent boundary count
    assert L.anuga_create(C.byref(desc), C.byref(p), C.byref(h)) == _native.ANUGA_E_ARG
    assert b"number_of_boundaries" in L.anuga_last_error(None)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_domain_fails_loudly_without_a_device():
    from anuga_core_amd import rectangular_cross_domain
    with pytest.raises(_native.DeviceError):
        rectangular_cross_domain(3, 3)


def test_ensemble_create_rejects_bad_sizes_before_touching_the_device():
    L = _native.lib()
    h = C.c_void_p()
    p = _native.AnugaParams(9.8, 1e-12, 1e-5, 1, 0, 1, 0, 1, 0, 0, 1, 2, 0, 1)
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int64))
    n = np.array([100, 0, 50], np.int64)
    m = np.array([10, 4, 6], np.int64)
    assert L.anuga_ensemble_create(3, ip(n), ip(m), C.byref(p), C.byref(h)) == _native.ANUGA_E_ARG
    assert b"number_of_elements" in L.anuga_last_error(None)
    assert L.anuga_ensemble_create(0, ip(n), ip(m), C.byref(p), C.byref(h)) == _native.ANUGA_E_ARG
    n = np.array([1 << 28, 1 << 28, 10], np.int64)   # too many triangles for one handle
    assert L.anuga_ensemble_create(3, ip(n), ip(m), C.byref(p), C.byref(h)) == _native.ANUGA_E_ARG
    assert not h.value


def test_build_matches_the_tuned_configuration():
    """The loaded library was compiled with the autotuner's recorded choice
    (anuga_core_amd/tuned.json), which build() reads."""
    from anuga_core_amd.build import tuned_config
    b, w = C.c_int64(), C.c_int64()
    assert _native.lib().anuga_build_config(C.byref(b), C.byref(w)) == 0
    cfg = tuned_config()
    assert (b.value, w.value) == (cfg["DE_BLOCK"], cfg["DE_MIXED_WAVES"])


def test_autotuner_choice_rule():
    """The tuner picks the block by the FP64 time, then the wave target by the
    mixed time at that block (here the recorded sweep gives 256 and 5)."""
    import json
    from anuga_core_amd.autotune import choose
    rows = [dict(DE_BLOCK=b, DE_MIXED_WAVES=w, precision=p, ms_per_step=t)
            for b, w, p, t in [(128, 5, "fp64", 0.9), (256, 5, "fp64", 0.8), (512, 5, "fp64", 0.85),
                               (128, 4, "mixed", 0.5), (256, 4, "mixed", 0.6), (256, 5, "mixed", 0.55),
                               (512, 4, "mixed", 0.4)]]
    assert choose(rows) == {"DE_BLOCK": 256, "DE_MIXED_WAVES": 5}
    tuned = os.path.join(os.path.dirname(_native.__file__), "tuned.json")
    with open(tuned) as f:
        rec = json.load(f)
    assert choose(rec["measured"]) == rec["config"]

"""The device's pow (anuga_core_amd/csrc/de_pow.h), compiled for the host with
g++, against the C library's pow that the reference calls
(sw_domain_orig.c:2827 friction, :710 weir): bit for bit on millions of
inputs over the ranges friction and the weir see, on the special values, and
within the method's 0.52 ulp bound of 80-digit decimal x^y. The tables are
regenerated from their recipe and compared with the header's (CPU only)."""

import ctypes as C
import math
import os
import re
import subprocess
import sys
import tempfile
from decimal import Decimal, getcontext

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "anuga_core_amd", "csrc")

SRC = r'''
#include "de_pow.h"
extern "C" void pow_many(long n, const double *x, double y, double *out) {
    for (long i = 0; i < n; i++) out[i] = de_pow_libm(x[i], y);
}
extern "C" void pow73_many(long n, const double *x, double *out) {
    for (long i = 0; i < n; i++) out[i] = de_pow_7_3(x[i]);
}
extern "C" void pow0385_many(long n, const double *x, double *out) {
    for (long i = 0; i < n; i++) out[i] = de_pow_0385(x[i]);
}
/* the C library's pow on the same inputs, and the count of differing bits */
extern "C" long count_mismatch(long n, const double *x, double y, const double *out) {
    long bad = 0;
    for (long i = 0; i < n; i++) {
        volatile double r = pow(x[i], y);
        if (memcmp((const void *)&r, &out[i], 8) != 0) bad++;
    }
    return bad;
}
'''

DP = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def lib():
    tmp = tempfile.mkdtemp()
    src = os.path.join(tmp, "p.cpp")
    so = os.path.join(tmp, "libp.so")
    open(src, "w").write(SRC)
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-shared", "-fPIC",
                    "-I" + CSRC, "-o", so, src], check=True)
    L = C.CDLL(so)
    L.pow_many.argtypes = [C.c_long, DP, C.c_double, DP]
    L.pow73_many.argtypes = [C.c_long, DP, DP]
    L.pow0385_many.argtypes = [C.c_long, DP, DP]
    L.count_mismatch.argtypes = [C.c_long, DP, C.c_double, DP]
    L.count_mismatch.restype = C.c_long
    return L


def _run(L, fn, x, *args):
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty_like(x)
    getattr(L, fn)(x.size, x.ctypes.data_as(DP), *args, out.ctypes.data_as(DP))
    return out


def _mismatch(L, x, y, out):
    return L.count_mismatch(x.size, x.ctypes.data_as(DP), y, out.ctypes.data_as(DP))


@pytest.mark.parametrize("y,lo,hi", [(7.0 / 3.0, 1e-12, 1e5), (7.0 / 3.0, 0.999, 1.001),
                                     (0.385, 1e-300, 1.0), (0.385, 0.9999, 1.0)])
def test_bitwise_equal_to_c_library_pow(lib, y, lo, hi):
    """4M log-uniform inputs per range, bit for bit. A restatement that drops
    the FMA contraction the library's build makes differs on about 1 in 1400."""
    rng = np.random.default_rng(11)
    x = np.exp(rng.uniform(math.log(lo), math.log(hi), 4_000_000))
    out = _run(lib, "pow_many", x, y)
    assert _mismatch(lib, x, y, out) == 0


def test_named_entry_points(lib):
    rng = np.random.default_rng(3)
    h = 10 ** rng.uniform(-5, 3, 100000)
    assert np.array_equal(_run(lib, "pow73_many", h), np.array([v ** (7.0 / 3.0) for v in h]))
    r = np.concatenate([[0.0, 1.0], rng.uniform(0, 1, 100000)])
    assert np.array_equal(_run(lib, "pow0385_many", r), np.array([v ** 0.385 for v in r]))


def test_within_published_bound_of_exact(lib):
    """|pow - x^y| <= 0.52 ulp against 80-digit decimal arithmetic."""
    getcontext().prec = 80
    e = Decimal(7.0 / 3.0)
    rng = np.random.default_rng(23)
    xs = np.concatenate([10 ** rng.uniform(-12, 4.5, 400), [1.0, 2.0, 0.5, 1e-5, 1e-12, 8.0]])
    ys = _run(lib, "pow73_many", xs)
    for x, r in zip(xs, ys):
        exact = (e * Decimal(float(x)).ln()).exp()
        assert abs(Decimal(float(r)) - exact) <= Decimal("0.52") * Decimal(math.ulp(float(r))), x


def test_special_values(lib):
    """Outside the method's domain (a run that has blown up: inf, NaN; zero,
    negative, subnormal; results beyond e^512) the library pow's IEEE special
    values come back instead of a table lookup with a garbage index."""
    xs = np.array([np.inf, np.nan, 0.0, -0.0, -1.0, -np.inf, 5e-324, 1e-310,
                   np.finfo(np.float64).max, 1e250, 1e-250, 1e-130])
    for y in (7.0 / 3.0, 0.385):
        out = _run(lib, "pow_many", xs, y)
        assert _mismatch(lib, xs, y, out) == 0
    y73 = _run(lib, "pow73_many", xs)
    assert np.isnan(y73[1]) and np.isnan(y73[4]) and y73[5] == np.inf
    assert y73[6] == 0.0 and y73[8] == np.inf and y73[9] == np.inf and y73[10] == 0.0


def test_tables_follow_the_recipe():
    """The header's tables are what gen_pow_tables.py derives from the
    method's recipe (decimal arithmetic), and that derivation reproduces the
    first log-table row and exp-table rows of the published tables."""
    sys.path.insert(0, CSRC)
    try:
        import gen_pow_tables as g
    finally:
        sys.path.remove(CSRC)
    text = open(os.path.join(CSRC, "de_pow.h")).read()
    body = text[text.index("kDePowLog[128 * 3]"):text.index("DE_POW_FN uint64_t de_pow_bits")]
    log_part, exp_part = body.split("kDePowExp[256]")
    log_vals = [float.fromhex(t) for t in re.findall(r"-?0x[0-9a-f.]+p[-+]\d+", log_part)]
    exp_vals = [int(t, 16) for t in re.findall(r"0x([0-9a-f]{16})ULL", exp_part)]
    assert log_vals == [v for row in g.log_table() for v in row]
    assert exp_vals == [v for pair in g.exp_table() for v in pair]
    assert g.log_table()[0] == (float.fromhex("0x1.6ap+0"), float.fromhex("-0x1.62c82f2b9c8p-2"),
                                float.fromhex("0x1.ab42428375680p-48"))
    assert g.exp_table()[1] == (0x3c9b3b4f1a88bf6e, 0x3feff63da9fb3335)

"""The product's correctly rounded log/exp (mc-mpi_amd/csrc/mc_math.h),
compiled for the host with gcc, against the oracle's independent correctly
rounded evaluation (x87 extended + binary128).  The same source is compiled
into the HIP kernels; tests/test_parity_gpu.py repeats the check on the
device."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SHIM = r"""
#include <stdint.h>
#include "mc_math.h"
void pm_log(const double *x, long n, double *y) { for (long i = 0; i < n; ++i) y[i] = mc_log(x[i]); }
void pm_exp(const double *x, long n, double *y) { for (long i = 0; i < n; ++i) y[i] = mc_exp(x[i]); }
void pm_exp_row(const double *x, long n, double *y) {
  const mc_coef K = mc_coef_default();
  for (long i = 0; i < n; ++i) {
    const mc_exp_row R = mc_exp_lookup(x[i]);
    y[i] = mc_exp_k_row(x[i], &R, &K);
  }
}
"""


@pytest.fixture(scope="module")
def pm(tmp_path_factory):
    d = tmp_path_factory.mktemp("pm")
    src = d / "pm.c"
    src.write_text(SHIM)
    so = d / "libpm.so"
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-fPIC", "-shared", "-ffp-contract=off",
                    "-I", os.path.join(ROOT, "mc-mpi_amd", "csrc"), "-o", str(so), str(src), "-lm"],
                   check=True)
    L = C.CDLL(str(so))
    for f in (L.pm_log, L.pm_exp, L.pm_exp_row):
        f.argtypes = [C.c_void_p, C.c_long, C.c_void_p]

    def run(which, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty_like(x)
        getattr(L, "pm_" + which)(x.ctypes.data, len(x), y.ctypes.data)
        return y
    return run


def _same(a, b):
    return (a.view(np.uint64) == b.view(np.uint64)) | (np.isnan(a) & np.isnan(b))


def test_log_random_rnd_real_outputs(pm, oracle):
    """log is applied to rnd_real outputs h = (double)s * 2^-63."""
    rng = np.random.default_rng(1)
    s = rng.integers(0, 2 ** 63, 400_000, dtype=np.uint64)
    s >>= rng.integers(0, 60, 400_000).astype(np.uint64)
    h = s.astype(np.float64) * 2.0 ** -63
    h = h[h > 0]
    assert np.all(_same(pm("log", h), oracle.math_eval("log", h)))


def test_log_near_one_and_structured(pm, oracle):
    k = np.arange(1, 100_001, dtype=np.float64)
    x = np.concatenate([1.0 + k * 2.0 ** -52, 1.0 - k * 2.0 ** -53,
                        1.0 + (np.random.default_rng(2).random(100_000) - 0.5) * 2.0 ** -8])
    assert np.all(_same(pm("log", x), oracle.math_eval("log", x)))


def test_log_full_range_and_specials(pm, oracle):
    rng = np.random.default_rng(3)
    bits = rng.integers(0, 2 ** 63, 200_000, dtype=np.uint64)
    x = bits.view(np.float64)
    x = np.concatenate([x, [0.0, -0.0, 1.0, np.inf, -1.0, np.nan, 5e-324, 2.2250738585072014e-308,
                            1.7976931348623157e308]])
    assert np.all(_same(pm("log", x), oracle.math_eval("log", x)))


def test_exp_transport_range(pm, oracle):
    """exp is applied to -sig_a * di <= 0 of every magnitude."""
    rng = np.random.default_rng(4)
    x = -rng.random(400_000) * 10.0 ** rng.uniform(-18, 2.9, 400_000)
    assert np.all(_same(pm("exp", x), oracle.math_eval("exp", x)))


def test_exp_structured_tiny_and_subnormal(pm, oracle):
    k = np.arange(1, 20_001, dtype=np.float64)
    x = np.concatenate([k * 2.0 ** -53, -k * 2.0 ** -53, k * 2.0 ** -60, -k * 2.0 ** -40,
                        -708.0 - np.random.default_rng(5).random(50_000) * 38.0,
                        [0.0, -0.0, np.inf, -np.inf, np.nan, -745.1332191019411, -745.14,
                         709.782712893384, 709.79, 2.0 ** -54, -(2.0 ** -54)]])
    assert np.all(_same(pm("exp", x), oracle.math_eval("exp", x)))


def test_exp_full_range(pm, oracle):
    x = (np.random.default_rng(6).random(300_000) - 0.5) * 1500.0
    assert np.all(_same(pm("exp", x), oracle.math_eval("exp", x)))


def test_exp_near_multiples_of_ln2_over_128(pm, oracle):
    """Arguments next to k*ln2/128 and k*ln2/1024 (the accurate path's and the
    fast path's reduction steps), where the reduced argument r nearly cancels
    and the two-fma reduction's error bound is tightest."""
    rng = np.random.default_rng(7)
    k = rng.integers(-130_000, 131_000, 300_000).astype(np.float64)
    off = (rng.integers(-1024, 1024, 300_000).astype(np.float64)
           * 2.0 ** -rng.integers(40, 70, 300_000).astype(np.float64))
    k10 = rng.integers(-1_040_000, 1_048_000, 300_000).astype(np.float64)
    x = np.concatenate([k * (np.log(2.0) / 128.0), k10 * (np.log(2.0) / 1024.0)]) \
        + np.concatenate([off, off])
    x = x[(x > -708.3) & (x < 709.7)]
    assert np.all(_same(pm("exp", x), oracle.math_eval("exp", x)))


def test_exp_in_two_halves_equals_exp(pm, oracle):
    """mc_exp_lookup + mc_exp_k_row (the row loaded a step ahead, as the
    event kernel does) give mc_exp's result bit for bit, and so the correctly
    rounded one, on every kind of argument the tests above use."""
    rng = np.random.default_rng(8)
    k = np.arange(1, 20_001, dtype=np.float64)
    x = np.concatenate([-rng.random(300_000) * 10.0 ** rng.uniform(-18, 2.9, 300_000),
                        (rng.random(200_000) - 0.5) * 1500.0, k * 2.0 ** -53, -k * 2.0 ** -60,
                        rng.integers(-1_040_000, 1_048_000, 100_000) * (np.log(2.0) / 1024.0),
                        [0.0, -0.0, np.inf, -np.inf, np.nan, -745.14, 709.79, -708.3, 708.3,
                         2.0 ** -54, 5e-324, -5e-324]])
    got = pm("exp_row", x)
    assert np.all(_same(got, pm("exp", x)))
    assert np.all(_same(got, oracle.math_eval("exp", x)))


def test_log_rows_next_to_one(pm, oracle):
    """The rows with the largest fast-path error bound are the buckets next to
    x = 1 (|log x| small); sample them densely, with every exponent."""
    rng = np.random.default_rng(8)
    m = 1.0 + (rng.random(300_000) - 0.5) * 2.0 ** -5
    x = np.ldexp(m, rng.integers(-3, 4, 300_000))
    assert np.all(_same(pm("log", x), oracle.math_eval("log", x)))


def test_math_tables_regenerate_identically():
    """gen_math_tables.py recomputes every table entry from its definition and
    asserts, per log-table row, the exactness of r = m*c - 1, the FastTwoSum
    ordering and the fast path's error bound (below 2^-68); its output must
    be the committed header byte for byte."""
    gen = os.path.join(ROOT, "mc-mpi_amd", "csrc", "gen_math_tables.py")
    out = subprocess.run(["python3", gen], check=True, capture_output=True, text=True).stdout
    with open(os.path.join(ROOT, "mc-mpi_amd", "csrc", "mc_math_tables.h")) as f:
        assert out == f.read()


CROSS_SHIM = r"""
#include <math.h>
#include <stdint.h>
#include "mc_math.h"
/* mode 0: the transport step's certificate, 1 - h > sig_i * di_edge * (1 + 2^-40);
   mode 1: a planted wrong-direction bound, (1 - h) / h in place of 1 - h
   (an upper bound of -log h, not a lower one). */
void pm_cross(const double *h, const double *sig_i, const double *di_edge, long n, int mode,
              int8_t *sure, int8_t *collided) {
  for (long i = 0; i < n; ++i) {
    const double reach = sig_i[i] * di_edge[i];
    const double lower = mode == 0 ? 1.0 - h[i] : (1.0 - h[i]) / h[i];
    sure[i] = !(sig_i[i] > 1e-4) || lower > fma(reach, 0x1p-40, reach);
    const double dq = sig_i[i] > 1e-4 ? -mc_log(h[i]) / sig_i[i] : 1.7976931348623157e308;
    collided[i] = dq < di_edge[i];
  }
}
"""


def test_crossing_certificate(tmp_path):
    """The transport step skips the log when 1 - h > sig_i * di_edge *
    (1 + 2^-40): since -log h >= 1 - h, the particle then crosses.  The
    shortcut must never call a collision a crossing, checked against the
    reference's comparison (correctly rounded -log(h) / sig_i < di_edge) on
    random draws, on h next to 1 with di_edge placed at the boundary of the
    test (ratios 1 +- 2^-k), and on the special inputs (h = 0 and 1, sig_i
    at and below EPS, di_edge of 0, -0, negative and MAXREAL).  A planted
    wrong-direction bound must fail the same check."""
    src = tmp_path / "pc.c"
    src.write_text(CROSS_SHIM)
    so = tmp_path / "libpc.so"
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-fPIC", "-shared", "-ffp-contract=off",
                    "-I", os.path.join(ROOT, "mc-mpi_amd", "csrc"), "-o", str(so), str(src), "-lm"],
                   check=True)
    L = C.CDLL(str(so))
    L.pm_cross.argtypes = [C.c_void_p] * 3 + [C.c_long, C.c_int, C.c_void_p, C.c_void_p]
    rng = np.random.default_rng(11)
    n = 300_000
    s = rng.integers(0, 2 ** 63, n, dtype=np.uint64)
    h_rand = s.astype(np.float64) * 2.0 ** -63
    h_near = 1.0 - rng.integers(1, 2 ** 24, n).astype(np.float64) * 2.0 ** -53
    h_mid = rng.random(n)
    hs, sigs, dis = [], [], []
    for h in (h_rand, h_near, h_mid):
        sig = np.exp(rng.normal(0.0, 1.5, n))
        k = rng.integers(8, 62, n).astype(np.float64)
        ratio = 1.0 + rng.choice([-1.0, 1.0], n) * np.ldexp(rng.random(n), -k.astype(int))
        ratio = np.where(rng.random(n) < 0.3, rng.random(n) * 3.0, ratio)
        hs.append(h)
        sigs.append(sig)
        dis.append(ratio * (1.0 - h) / sig)
        # di_edge at -log(h) / sig_i: inside the band where a wrong-direction
        # bound claims a crossing
        hs.append(h)
        sigs.append(sig)
        dis.append(-np.log(np.maximum(h, 1e-300)) / sig * (1.0 + 2.0 ** -20))
    special_h = np.array([0.0, 1.0, 2.0 ** -63, 1.0 - 2.0 ** -53, 0.5])
    special_sig = np.array([0.0, 1e-4, 5e-5, np.nextafter(1e-4, 1.0), 1.0, 1e300])
    special_di = np.array([0.0, -0.0, -1e-3, 1.7976931348623157e308, 1e-300, 1.0])
    g = np.array(np.meshgrid(special_h, special_sig, special_di)).reshape(3, -1)
    hs.append(g[0])
    sigs.append(g[1])
    dis.append(g[2])
    h = np.ascontiguousarray(np.concatenate(hs))
    sig = np.ascontiguousarray(np.concatenate(sigs))
    di = np.ascontiguousarray(np.concatenate(dis))
    sure = np.empty(len(h), np.int8)
    coll = np.empty(len(h), np.int8)

    def run(mode):
        L.pm_cross(h.ctypes.data, sig.ctypes.data, di.ctypes.data, len(h), mode,
                   sure.ctypes.data, coll.ctypes.data)
        return sure.astype(bool), coll.astype(bool)

    sure0, coll0 = run(0)
    assert not np.any(sure0 & coll0)
    assert sure0.mean() > 0.1 and coll0.mean() > 0.1      # both outcomes exercised
    sure1, coll1 = run(1)
    assert np.count_nonzero(sure1 & coll1) > 1000           # the planted bound is caught


DIV_SHIM = r"""
#include <stdint.h>
#include "mc_math.h"
void pm_div_by_rcp(const double *a, const double *b, long n, double *q) {
  for (long i = 0; i < n; ++i) q[i] = mc_div_by_rcp(a[i], b[i], 1.0 / b[i]);
}
"""


def test_div_by_reciprocal_equals_ieee_division(tmp_path):
    """The edge distance (x_new_edge - x) / mu is formed from a stored
    RN(1/mu) with two fma corrections; it must equal the IEEE quotient bit
    for bit, including divisors with all-ones significands, operands next to
    powers of two and the fallback range (zero numerators, extreme ratios)."""
    src = tmp_path / "pd.c"
    src.write_text(DIV_SHIM)
    so = tmp_path / "libpd.so"
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-fPIC", "-shared", "-ffp-contract=off",
                    "-I", os.path.join(ROOT, "mc-mpi_amd", "csrc"), "-o", str(so), str(src), "-lm"],
                   check=True)
    L = C.CDLL(str(so))
    L.pm_div_by_rcp.argtypes = [C.c_void_p, C.c_void_p, C.c_long, C.c_void_p]
    rng = np.random.default_rng(10)
    n = 400_000
    a = [rng.random(n) * 2 - 1,
         np.ldexp(1 + rng.random(n), rng.integers(-100, 100, n)),
         (rng.integers(-10000, 10001, n) * 1e-4 - rng.random(n)),
         (np.uint64(0x3FF0000000000000) + rng.integers(0, 1 << 20, n, dtype=np.uint64)).view(np.float64),
         np.concatenate([np.zeros(1000), np.ldexp(1.0, rng.integers(-1000, 1000, n - 1000))])]
    b = [rng.random(n) * 2 - 1,
         np.ldexp(1 + rng.random(n), rng.integers(-60, 60, n)),
         np.sqrt(rng.random(n)),
         (np.uint64(0x3FEFFFFFFFFFFFFF) - rng.integers(0, 1 << 16, n, dtype=np.uint64)).view(np.float64),
         np.ldexp(1 + rng.random(n), rng.integers(-60, 60, n))]
    a = np.concatenate(a)
    b = np.concatenate(b) * np.where(rng.random(5 * n) < 0.5, -1.0, 1.0)
    keep = np.abs(b) >= 2.0 ** -60
    a, b = np.ascontiguousarray(a[keep]), np.ascontiguousarray(b[keep])
    q = np.empty_like(a)
    L.pm_div_by_rcp(a.ctypes.data, b.ctypes.data, len(a), q.ctypes.data)
    with np.errstate(all="ignore"):
        assert np.all(_same(q, a / b))

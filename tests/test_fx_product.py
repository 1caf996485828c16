"""The product's fixed-point tally arithmetic (mc-mpi_amd/csrc/mc_fx.h),
compiled for the host with gcc, against the oracle's independent 128-bit
implementation: term -> (lo, hi, flags) for every exponent and sign
(including the boundaries of the kernels' fast path, powers of two whose low
word is zero, zeros, subnormals, out-of-range and non-finite terms), and the
correctly rounded sum -> double conversion.  The same source is compiled
into every HIP kernel."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SHIM = r"""
#include <stdint.h>
#include "mc_fx.h"
void px_convert(const double *v, long n, uint64_t *out, uint32_t *flags, int positive_only) {
  for (long i = 0; i < n; ++i) {
    uint32_t f = 0;
    mc_fx r = mc_fx_from_double_p(v[i], &f, positive_only);
    out[2 * i] = r.lo;
    out[2 * i + 1] = r.hi;
    flags[i] = f;
  }
}
void px_finalize(const uint64_t *fx, const uint32_t *flags, long n, double *out) {
  for (long i = 0; i < n; ++i) out[i] = mc_fx_finalize(fx[2 * i], fx[2 * i + 1], flags[i]);
}
"""


@pytest.fixture(scope="module")
def px(tmp_path_factory):
    d = tmp_path_factory.mktemp("px")
    src = d / "px.c"
    src.write_text(SHIM)
    so = d / "libpx.so"
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-fPIC", "-shared", "-ffp-contract=off",
                    "-I", os.path.join(ROOT, "mc-mpi_amd", "csrc"), "-o", str(so), str(src), "-lm"],
                   check=True)
    L = C.CDLL(str(so))
    L.px_convert.argtypes = [C.c_void_p, C.c_long, C.c_void_p, C.c_void_p, C.c_int]
    L.px_finalize.argtypes = [C.c_void_p, C.c_void_p, C.c_long, C.c_void_p]
    return L


def _terms():
    rng = np.random.default_rng(17)
    e = np.arange(-1080, 60)
    pow2 = np.ldexp(1.0, e)
    edge = np.concatenate([pow2, -pow2, pow2 * (1 + 2.0 ** -52), -pow2 * (1 - 2.0 ** -53),
                           np.nextafter(pow2, 0), -np.nextafter(pow2, np.inf)])
    rnd = rng.uniform(1, 2, 400_000) * np.ldexp(1.0, rng.integers(-100, 45, 400_000))
    rnd *= np.where(rng.random(400_000) < 0.5, -1.0, 1.0)
    special = np.array([0.0, -0.0, 5e-324, -5e-324, 2.0 ** -89, 3 * 2.0 ** -90, -3 * 2.0 ** -90,
                        2.0 ** 39, -2.0 ** 39, np.nextafter(2.0 ** 39, 0), np.inf, -np.inf, np.nan])
    return np.concatenate([edge, rnd, special])


@pytest.mark.parametrize("positive_only", [0, 1])
def test_term_conversion_matches_oracle(px, oracle, positive_only):
    """Both sign variants of the fast path (signed: adjoint terms; positive
    only: transport deposits) give the oracle's integers for every term."""
    v = _terms()
    out = np.zeros(2 * len(v), np.uint64)
    fl = np.zeros(len(v), np.uint32)
    px.px_convert(v.ctypes.data, len(v), out.ctypes.data, fl.ctypes.data, positive_only)
    want, wfl = oracle.fx_convert(v)
    bad = np.flatnonzero((out[0::2] != want[0::2]) | (out[1::2] != want[1::2]) | (fl != wfl))
    assert len(bad) == 0, v[bad[:5]]


def test_sum_conversion_matches_oracle(px, oracle):
    rng = np.random.default_rng(18)
    n = 300_000
    fx = rng.integers(0, 2 ** 63, 2 * n, dtype=np.uint64) * np.uint64(2) + \
        rng.integers(0, 2, 2 * n, dtype=np.uint64)
    fx[1::2] >>= rng.integers(0, 64, n).astype(np.uint64)       # every magnitude of the high word
    fx[1:2000:2] = 0                                              # high word zero
    fx[3000:6000:2] = 0                                           # low word zero
    fl = np.zeros(n, np.uint32)
    fl[-8:] = [1, 2, 4, 8, 6, 3, 5, 0]
    got = np.zeros(n)
    px.px_finalize(fx.ctypes.data, fl.ctypes.data, n, got.ctypes.data)
    want = oracle.fx_finalize(fx, fl)
    same = (got.view(np.uint64) == want.view(np.uint64)) | (np.isnan(got) & np.isnan(want))
    assert same.all()

#!/usr/bin/env python3
"""Generate mc_math_tables.h: the constant tables of the correctly rounded
FP64 log/exp used by the transport kernels.

Every value is computed here from its mathematical definition with Python's
`decimal` module at 80 significant digits and split into double-double
(hi + lo) pieces with correctly rounded conversions, so the tables carry
about 106 bits.  Nothing is taken from any existing libm.

log table (Tang-style reduction).  With tmp = bits(x) - OFF, x = 2^e * m
where e = tmp >> 52 (arithmetic) and m in [0.7075, 1.4150); the index
j = bits 43..51 of tmp (0..511) cuts that range into 512 buckets of equal
width in the bit pattern (2^-10 in m below 1, 2^-9 above).  OFF is chosen
so that m = 1 sits at the centre of its bucket, whose c is 1 (r = m - 1,
|r| < 2^-10, so log x near 1 keeps its relative accuracy).  Every other
bucket has c_j, a 9-significant-bit approximation of 1/m over the bucket,
so that r = m*c_j - 1 is exact in one fma.  The table stores c_j and
-log(c_j) as head + mid + lo, the head on the 2^-42 grid so that e*ln2_hi +
head is exact.  |r|, the exactness of r and the FastTwoSum ordering of head
and r are verified below with exact rational arithmetic over every bucket.

exp table: 2^(j/128), j = 0..127, as double-double.

Run:  python3 gen_math_tables.py > mc_math_tables.h
"""
import decimal
from decimal import Decimal
from fractions import Fraction
import math
import struct

decimal.getcontext().prec = 80

LOG_BITS = 9
# near the bits of sqrt(1/2); bits(1.0) - OFF = 299 * 2^43 + 2^42, so m = 1
# is the centre of bucket 299
OFF = 0x3FE6A40000000000
ONE_BITS = 0x3FF0000000000000
J_ONE = (ONE_BITS - OFF) >> 43
assert ((ONE_BITS - OFF) & ((1 << 43) - 1)) == 1 << 42


def f64_bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def bits_f64(b):
    return struct.unpack("<d", struct.pack("<Q", b))[0]


def dd(value):
    """Split a Decimal into (hi, lo) doubles, each correctly rounded."""
    hi = float(value)
    lo = float(value - Decimal(hi))
    return hi, lo


def to_nbits(q, nbits):
    """Round a positive Fraction to nbits significant bits (nearest)."""
    e = math.floor(math.log2(q))
    # guard against log2 rounding at exact powers of two
    while Fraction(2) ** e > q:
        e -= 1
    while Fraction(2) ** (e + 1) <= q:
        e += 1
    scale = Fraction(2) ** (nbits - 1 - e)
    k = round(q * scale)
    return Fraction(k) / scale


def ulp_of(x):
    """ulp of a positive double given as Fraction (normal range)."""
    e = math.floor(math.log2(x))
    while Fraction(2) ** e > x:
        e -= 1
    while Fraction(2) ** (e + 1) <= x:
        e += 1
    return Fraction(2) ** (e - 52)


def log_fast_error(a, b, c, head, mid, rmax):
    """Bound on the relative error of mc_log's fast path (the double-double
    hi + lo before the rounding test) over m in [a, b], e = 0 (for e != 0
    the result exceeds 0.34 in magnitude and the bound is far smaller).

    Absolute error terms, u = 2^-53, R = max |r|, |P| <= 0.34:
      t = (r2h*r)*p: three roundings and the error of p's five fmas, with
        the r2l*r term left out: below 1.7 u R^3 (counted as 2 u R^3);
      fma(-0.5, r2l, t), serr3 + ..., lo1 + ...: three roundings of sums
        led by t, below 1.1 u R^3 plus u(|mid| + u(|head| + 2R));
      lo0 + serr2 and the table's dropped third part: below 2 u |mid|;
      the series tail r^9/9 - ...: below R^9 / (9 (1 - R)).
    Relative to min |log m| over the bucket; in the bucket of m = 1 (c = 1,
    head = mid = 0) the terms scale with |r|^3 and |log m| >= |r| (1 - R/2),
    so the bound is taken per unit |r|."""
    u = Fraction(1, 2 ** 53)
    R = rmax
    if head == 0 and mid == 0:
        num = Fraction(31, 10) * u * R * R + 6 * u * u + R ** 8 / (9 * (1 - R))
        return num / (1 - R / 2)
    absd = (Fraction(31, 10) * u * R ** 3 + 3 * u * abs(Fraction(mid))
            + u * u * (abs(Fraction(head)) + 2 * R) + R ** 9 / (9 * (1 - R)))
    lmin = min(abs(Decimal(a.numerator) / Decimal(a.denominator)).ln().copy_abs(),
               abs(Decimal(b.numerator) / Decimal(b.denominator)).ln().copy_abs())
    return absd / Fraction(lmin)


def log_table():
    rows = []
    rmax_all = Fraction(0)
    err_all = Fraction(0)
    for j in range(1 << LOG_BITS):
        # closed interval of m whose bit pattern lies in bucket j (e = 0)
        a = Fraction(bits_f64(OFF + (j << 43)))
        b = Fraction(bits_f64(OFF + ((j + 1) << 43) - 1))
        if j == J_ONE:
            assert a < 1 < b
            cands = [Fraction(1)]
        else:
            assert (a >= 1) == (b >= 1), j  # no other bucket straddles m = 1
            c0 = to_nbits(2 / (a + b), 9)
            ce = math.floor(math.log2(c0))
            while Fraction(2) ** ce > c0:
                ce -= 1
            ulp_c = Fraction(2) ** (ce - 8)
            cands = [c0, c0 - ulp_c, c0 + ulp_c, Fraction(1)]
        best = None
        for c in cands:
            rmax = max(abs(a * c - 1), abs(b * c - 1))
            if rmax >= Fraction(1, 362):  # 1/362 < 2^-8.49: series tail bound
                continue
            cd = float(c)
            assert Fraction(cd) == c
            L = -(Decimal(cd).ln())
            # head on the 2^-42 grid (so e*LN2_HI + head is exact for
            # |e| < 2^11, LN2_HI having 42 bits), then two double tails
            head = Fraction(round(Fraction(L) * 2 ** 42), 2 ** 42)
            hi = float(head)
            assert Fraction(hi) == head
            rest = L - Decimal(hi)
            mid = float(rest)
            lo = float(rest - Decimal(mid))
            # with e = 0 the sum hi + r is formed by FastTwoSum: it needs
            # the exponent of hi to be at least that of r, or hi = 0
            if hi != 0 and math.floor(math.log2(abs(hi))) < math.floor(math.log2(float(rmax))):
                continue
            # exactness of r = m*c - 1: r is an integer multiple of
            # ulp(m) * ulp_9bit(c); it is exact if that integer is < 2^53
            # (c = 1 near m = 1 included: r = m - 1)
            ce = math.floor(math.log2(c))
            while Fraction(2) ** ce > c:
                ce -= 1
            ulp_c = Fraction(2) ** (ce - 8)
            assert (c / ulp_c).denominator == 1
            for m_end in (a, b):
                ulp_m = Fraction(2) ** -52 if m_end >= 1 else Fraction(2) ** -53
                if rmax / (ulp_m * ulp_c) >= 2 ** 53:
                    break
            else:
                err = log_fast_error(a, b, c, hi, mid, rmax)
                if best is None or err < best[0]:
                    best = (err, rmax, cd, hi, mid, lo)
        assert best is not None, j
        err_all = max(err_all, best[0])
        rmax_all = max(rmax_all, best[1])
        rows.append((j,) + best[2:])
    # mc_rounds_to_hi (mc_math.h) is stated for errors below 2^-68
    assert err_all < Fraction(1, 2 ** 68), math.log2(err_all)
    return float(rmax_all), float(err_all), rows


def exp_table(n):
    rows = []
    for j in range(n):
        v = Decimal(2) ** (Decimal(j) / Decimal(n))
        rows.append(dd(v))
    return rows


def split_nbits(value, nbits):
    """hi = value rounded to nbits significant bits (as double), then dd tail."""
    q = Fraction(value)  # Decimal -> exact Fraction
    hi = float(to_nbits(q, nbits))
    rest = value - Decimal(hi)
    mid = float(rest)
    lo = float(rest - Decimal(mid))
    return hi, mid, lo


def hexf(x):
    return float.hex(x)


def main():
    ln2 = Decimal(2).ln()
    out = []
    w = out.append
    w("/* Generated by gen_math_tables.py -- do not edit.  Values computed from")
    w("   their definitions with 80-digit decimal arithmetic. */")
    w("#ifndef MC_MATH_TABLES_H")
    w("#define MC_MATH_TABLES_H")
    w("")
    rmax, err, rows = log_table()
    w(f"#define MC_LOG_N {len(rows)}")
    w(f"#define MC_LOG_OFF 0x{OFF:016X}ULL")
    w(f"/* largest |r| over the table: 2^{math.log2(rmax):.3f}; bound on the relative")
    w(f"   error of the fast path before its rounding test: 2^{math.log2(err):.2f} */")
    w("/* {c_j, head, mid, lo} with -log(c_j) = head + mid + lo (head on the 2^-42")
    w("   grid), for bucket j = bits 43..51 of bits(x) - MC_LOG_OFF */")
    w("MC_TAB double mc_log_tab[MC_LOG_N][4] = {")
    for (i, c, hi, mid, lo) in rows:
        w(f"  {{{hexf(c)}, {hexf(hi)}, {hexf(mid)}, {hexf(lo)}}}, /* {i} */")
    w("};")
    w("")
    w("/* {hi, lo} of 2^(j/128) (accurate path) */")
    w("MC_TAB double mc_exp_tab[128][2] = {")
    for j, (hi, lo) in enumerate(exp_table(128)):
        w(f"  {{{hexf(hi)}, {hexf(lo)}}}, /* {j} */")
    w("};")
    w("")
    w("/* {hi, lo} of 2^(j/1024) (fast path) */")
    w("MC_TAB double mc_exp_tab10[1024][2] = {")
    for j, (hi, lo) in enumerate(exp_table(1024)):
        w(f"  {{{hexf(hi)}, {hexf(lo)}}}, /* {j} */")
    w("};")
    w("")
    # ln2 in three pieces; the first has 42 bits so e*LN2_HI is exact |e|<2^11
    h, m, l = split_nbits(ln2, 42)
    w(f"#define MC_LN2_HI {hexf(h)}")
    w(f"#define MC_LN2_MID {hexf(m)}")
    w(f"#define MC_LN2_LO {hexf(l)}")
    # ln2/128 in three pieces; the first has 32 bits so k*L1 is exact |k|<2^21
    h, m, l = split_nbits(ln2 / 128, 32)
    w(f"#define MC_LN2_128_1 {hexf(h)}")
    w(f"#define MC_LN2_128_2 {hexf(m)}")
    w(f"#define MC_LN2_128_3 {hexf(l)}")
    w(f"#define MC_INV_LN2_128 {hexf(float(Decimal(128) / ln2))}")
    # ln2/128 in two pieces for the fast path of exp: A has 36 significant
    # bits, so k*A is exact for |k| < 2^17 (|x| < 709.7 gives |k| <= 131057);
    # B = RN(ln2/128 - A) with |B| <= 2^-44 and |ln2/128 - A - B| <= 2^-97,
    # so |k| * that residual stays below 2^-80.
    a, b, _ = split_nbits(ln2 / 128, 36)
    assert Decimal(709.7) * 128 / ln2 < 2 ** 17
    assert abs(Decimal(b)) <= Decimal(2) ** -44
    assert abs(ln2 / 128 - Decimal(a) - Decimal(b)) <= Decimal(2) ** -97
    assert to_nbits(Fraction(a), 36) == Fraction(a)
    w(f"#define MC_LN2_128_A {hexf(a)}")
    w(f"#define MC_LN2_128_B {hexf(b)}")
    # ln2/1024 likewise for the fast path of exp: A has 33 significant bits,
    # so k*A is exact for |k| < 2^20 (|x| < 709.7 gives |k| <= 1048464);
    # |B| <= 2^-44 and |ln2/1024 - A - B| <= 2^-97, so |k| times that
    # residual stays below 2^-77.
    a, b, _ = split_nbits(ln2 / 1024, 33)
    assert Decimal(709.7) * 1024 / ln2 < 2 ** 20
    assert abs(Decimal(b)) <= Decimal(2) ** -44
    assert abs(ln2 / 1024 - Decimal(a) - Decimal(b)) <= Decimal(2) ** -97
    assert to_nbits(Fraction(a), 33) == Fraction(a)
    w(f"#define MC_LN2_1024_A {hexf(a)}")
    w(f"#define MC_LN2_1024_B {hexf(b)}")
    w(f"#define MC_INV_LN2_1024 {hexf(float(Decimal(1024) / ln2))}")
    w("")
    # 1/n (n = 1..16) and 1/n! (n = 0..16) as double-double
    w("/* {hi, lo} of 1/n, n = 0..16 (entry 0 unused) */")
    w("MC_TAB double mc_inv_n[17][2] = {")
    w("  {0x0p+0, 0x0p+0},")
    for n in range(1, 17):
        hi, lo = dd(Decimal(1) / Decimal(n))
        w(f"  {{{hexf(hi)}, {hexf(lo)}}}, /* 1/{n} */")
    w("};")
    w("/* {hi, lo} of 1/n!, n = 0..16 */")
    w("MC_TAB double mc_inv_fact[17][2] = {")
    for n in range(0, 17):
        hi, lo = dd(Decimal(1) / Decimal(math.factorial(n)))
        w(f"  {{{hexf(hi)}, {hexf(lo)}}}, /* 1/{n}! */")
    w("};")
    w("")
    w("#endif")
    print("\n".join(out))


if __name__ == "__main__":
    main()

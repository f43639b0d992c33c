"""The split-factor evaluations of the pair kernels against their stated error bounds.

The split factor of the short-range pair force, G(u) = erfc(u) + 2u/sqrt(pi) exp(-u^2),
u = r/2rs (src/fmm.c:845-847 evaluates it with erfc and exp), is read by the P2P
kernels from piecewise-polynomial tables (photons-2.0_amd/csrc/split_table.h, written
by split_table.py in the same directory); the FP32 M2L radial factors use erfc through a
fit of erfcx in t = 1/(1 + u/2) (fmm_math.h, from fit_erfcx.py).  The coefficients are
read from the sources and evaluated as the kernels evaluate them (Horner, one fused
multiply-add per step, rounded to the kernel's precision), against mpmath in 30- to
40-digit arithmetic.
"""
import os
import re
import sys

import mpmath as mp
import numpy as np
import pytest

SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "photons-2.0_amd", "csrc",
                   "fmm_math.h")
mp.mp.dps = 30


def body(name):
    text = open(SRC).read()
    i = text.index(name + "(")
    j = text.index("\n}\n", i)
    return text[i:j]


def float_coefficients(name):
    """highest power first: the initial value, then each fmaf's constant"""
    b = body(name)
    first = re.search(r"float q = ([-+0-9.e]+)f;", b).group(1)
    rest = re.findall(r"__builtin_fmaf\(q, t, ([-+0-9.e]+)f\)", b)
    return [np.float32(x) for x in [first] + rest]


def double_coefficients(name):
    """highest power first (the source lists them lowest first)"""
    b = body(name)
    nums = re.findall(r"[-+]?\d\.\d+e[-+]\d+", b[b.index("{"):])
    return [float(x) for x in reversed(nums)]


def horner(coef, t, dtype):
    q = dtype(coef[0])
    for c in coef[1:]:
        q = dtype(np.float64(q) * np.float64(t) + np.float64(c))
    return q


def erfcx(u):
    u = mp.mpf(float(u))
    return mp.exp(u * u) * mp.erfc(u)


def grid(u_hi, n):
    lin = np.linspace(0.0, min(u_hi, 8.0), n)
    if u_hi <= 8.0:
        return lin
    return np.concatenate([lin, np.geomspace(8.0, u_hi, n)])


def max_rel(coef, times_t, u_hi, dtype, n=1500):
    worst = 0.0
    for u in grid(u_hi, n):
        t = dtype(1.0 / (1.0 + 0.5 * u))
        p = horner(coef, t, dtype)
        if times_t:
            p = dtype(np.float64(t) * np.float64(p))
        worst = max(worst, float(abs(mp.mpf(float(p)) / erfcx(u) - 1)))
    return worst


def test_float_erfcx_tq_bound():
    """t * Q10(t): relative error <= 3.2e-7 in float over u in [0, 1e4]"""
    c = float_coefficients("erfcx_tq")
    assert len(c) == 11
    e = max_rel(c, True, 1e4, np.float32)
    print(f"erfcx_tq: {e:.2e}")
    assert e <= 3.2e-7


def test_double_erfcx_tq20_bound():
    """t * Q20(t), the FP64 pair force beyond its table: relative error <= 2.6e-14 in
    double over u >= 0 (checked up to 1e4, where t = 2e-4 and t Q20(t) follows the
    1/(u sqrt(pi)) tail)"""
    c = double_coefficients("erfcx_tq20")
    assert len(c) == 21
    e = max_rel(c, True, 1e4, np.float64)
    print(f"erfcx_tq20: {e:.2e}")
    assert e <= 2.65e-14  # 2.6e-14 to two digits


def test_a_wrong_coefficient_is_caught():
    """the check has teeth: the constant term perturbed by 1e-4 of itself fails"""
    c = float_coefficients("erfcx_tq")
    c[-1] = np.float32(c[-1] * (1 + 1e-4))
    assert max_rel(c, True, 6.0, np.float32, 300) > 3.2e-7


# ---------------------------------------------------------------- split-factor tables

TABLE_SRC = os.path.join(os.path.dirname(SRC), "split_table.h")


def table_rows(name):
    """the rows of a table in split_table.h, and its N and U"""
    text = open(TABLE_SRC).read()
    n = int(re.search(name + r"N = (\d+);", text).group(1))
    u = float(re.search(name + r"U = ([0-9.]+);", text).group(1))
    i = text.index("{", text.index(name + "["))
    j = text.index("};", i)
    rows = [[float(x.rstrip("f")) for x in re.findall(r"[-+]?\d\.\d+e[-+]\d+f?", r)]
            for r in text[i:j].split("},")[:-1]]
    return rows, n, u


def G(u):
    u = mp.mpf(u)
    return mp.erfc(u) + 2 * u / mp.sqrt(mp.pi) * mp.exp(-u * u)


def table_error(rows, n, umax, dtype, npts=8000):
    """largest |table - G| and largest relative error for u < 3, at the u the kernel
    indexes with (v = u N / U rounded to dtype, clamped to N), evaluated as the kernel
    does: i = floor(v), f = v - i, Horner with one rounding per step"""
    R = np.array(rows, dtype=dtype)
    worst, worst_rel = 0.0, 0.0
    with mp.workdps(40):
        for u in np.linspace(0.0, 1.1 * umax, npts):
            v = min(dtype(dtype(u) * dtype(n / umax)), dtype(n))
            i = int(v)
            f = dtype(v - dtype(i))
            c = R[i]
            g = c[-1]
            for k in range(len(c) - 2, -1, -1):
                g = dtype(np.longdouble(g) * np.longdouble(f) + np.longdouble(c[k]))
            ref = G(float(v) * umax / n) if v < n else G(u)
            e = abs(float(mp.mpf(float(g)) - ref))
            worst = max(worst, e)
            if u < 3.0:
                worst_rel = max(worst_rel, e / float(ref))
    return worst, worst_rel


def test_split_table_header_is_the_generators_output():
    """split_table.h is what split_table.py writes (no hand edits, no stale table)"""
    sys.path.insert(0, os.path.dirname(SRC))
    import split_table
    with mp.workdps(40):
        assert split_table.render() == open(TABLE_SRC).read()


def test_fp32_split_table_bound():
    """the FP32 P2P table (cubic, 256 intervals on u in [0, 9.5)): |error| <= 6.5e-8 (G <= 1),
    relative <= 6e-7 for u < 3, zero from u = 9.5 on (G(9.5) = 6e-39, below float's
    normal range)"""
    rows, n, umax = table_rows("kSplitTab32")
    assert (len(rows), len(rows[0])) == (n + 1, 4) and rows[n] == [0.0] * 4
    e, er = table_error(rows, n, umax, np.float32)
    print(f"FP32 table: max |error| {e:.2e}, relative (u < 3) {er:.2e}")
    assert e <= 6.5e-8 and er <= 6e-7


def test_fp64_split_table_bound():
    """the FP64 P2P table (degree 9, 48 intervals on u in [0, 6)): |error| <= 2e-15,
    relative <= 4e-15 for u < 3, zero from u = 6 on (G(6) = 1.6e-15)"""
    rows, n, umax = table_rows("kSplitTab64")
    assert (len(rows), len(rows[0])) == (n + 1, 10) and rows[n] == [0.0] * 10
    e, er = table_error(rows, n, umax, np.float64)
    print(f"FP64 table: max |error| {e:.2e}, relative (u < 3) {er:.2e}")
    assert e <= 2e-15 and er <= 4e-15


def test_a_wrong_table_entry_is_caught():
    """the check has teeth: one coefficient of one FP32 row off by 1e-6 fails"""
    rows, n, umax = table_rows("kSplitTab32")
    rows[20][1] += 1e-6
    e, _ = table_error(rows, n, umax, np.float32, 3000)
    assert e > 6.5e-8

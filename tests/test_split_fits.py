"""The erfcx fits of the pair kernels (photons-2.0_amd/csrc/fmm_math.h) against their
stated error bounds.

The split factor G(u) = exp(-u^2) (erfcx(u) + 2u/sqrt(pi)) of the short-range force
(src/fmm.c:834-852 evaluates it with erfc and exp) is computed on the GPU from
polynomial fits of erfcx in t = 1/(1 + u/2), written by photons-2.0_amd/csrc/
fit_erfcx.py.  The coefficients are read from the kernel source and evaluated as the
kernels evaluate them (Horner, one fused multiply-add per step, rounded to the kernel's
precision), against erfcx from mpmath in 30-digit arithmetic.
"""
import os
import re

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


def test_float_erfcx_t6_bound():
    """P6(t) of the FP32 P2P kernel: relative error <= 4.7e-6 in float over u in
    [0, 6]; beyond, the polynomial stays within [0, 0.1] (exp(-u^2) < 3e-16 there)"""
    c = float_coefficients("erfcx_t6")
    assert len(c) == 7
    e = max_rel(c, False, 6.0, np.float32, 3000)
    print(f"erfcx_t6: {e:.2e}")
    assert e <= 4.7e-6
    for u in np.geomspace(6.0, 1e4, 200):
        v = float(horner(c, np.float32(1.0 / (1.0 + 0.5 * u)), np.float32))
        assert 0.0 <= v <= 0.1


def test_double_erfcx_tq20_bound():
    """t * Q20(t) of the FP64 P2P kernel: relative error <= 2.6e-14 in double over u >= 0
    (checked up to 1e4, where t = 2e-4 and t Q20(t) follows the 1/(u sqrt(pi)) tail)"""
    c = double_coefficients("erfcx_tq20")
    assert len(c) == 21
    e = max_rel(c, True, 1e4, np.float64)
    print(f"erfcx_tq20: {e:.2e}")
    assert e <= 2.65e-14  # 2.6e-14 to two digits


@pytest.mark.parametrize("name", ["erfcx_tq", "erfcx_t6"])
def test_a_wrong_coefficient_is_caught(name):
    """the check has teeth: the constant term perturbed by 1e-4 of itself fails"""
    c = float_coefficients(name)
    c[-1] = np.float32(c[-1] * (1 + 1e-4))
    bound = 3.2e-7 if name == "erfcx_tq" else 4.7e-6
    assert max_rel(c, name == "erfcx_tq", 6.0, np.float32, 300) > bound

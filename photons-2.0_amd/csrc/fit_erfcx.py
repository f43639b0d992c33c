"""Fits of erfcx(u) = exp(u^2) erfc(u) used by kernels in fmm_math.h.

erfc(u) = exp(-u^2) erfcx(u); the kernels evaluate erfcx as t times a polynomial in
t = 1 / (1 + u/2) (the variable of the Numerical Recipes erfcc form, Press et al.,
sec. 6.2), which maps u in [0, inf) onto t in (0, 1]:

  erfcx_tq    float   t * Q10(t),  u in [0, 1e4]   (the FP32 M2L radial factors)
  erfcx_tq20  double  t * Q20(t),  u >= 0          (FP64 pair force beyond its table)

(The P2P kernels read the split factor of the pair force from the tables that
split_table.py writes, on u in [0, U); erfcx_tq20 gives the FP64 one from U on.)

Each fit minimises the relative error: least squares on Chebyshev nodes in t, weights
1/erfcx, then a few passes of iteratively reweighted least squares that push the
largest relative errors down (Lawson's algorithm), all in 50-digit arithmetic
(mpmath).  The float fits are then rounded to float.  Run

    python photons-2.0_amd/csrc/fit_erfcx.py

to print the coefficients (highest power first, as the Horner loops in fmm_math.h take
them) and their errors; it takes about two minutes.  It prints fits of the same form
with relative errors 1.8e-7 (erfcx_tq) and 3.7e-14 (erfcx_tq20).  The coefficients
fmm_math.h holds came from an earlier run of this procedure with other node sets and
iteration counts (3.2e-7, 2.6e-14); tests/test_split_fits.py checks them against those
bounds, evaluated as the kernels evaluate them.
"""
from __future__ import annotations

import mpmath as mp
import numpy as np

mp.mp.dps = 50


def erfcx_mp(u):
    u = mp.mpf(u)
    return mp.exp(u * u) * mp.erfc(u)


def t_of(u):
    return 1 / (1 + u / 2)


def u_of(t):
    return 2 * (1 / t - 1)


def fit(deg, times_t, t_lo, t_hi, nodes=600, passes=12):
    """relative-error fit of erfcx(u(t)) on t in [t_lo, t_hi]: coefficients c[0..deg]
    (lowest power first) of P(t), the model being P(t) or t * P(t)"""
    ts = [mp.mpf(t_lo) + (mp.mpf(t_hi) - t_lo) * (1 - mp.cos(mp.pi * (k + mp.mpf(0.5)) / nodes)) / 2
          for k in range(nodes)]
    f = [erfcx_mp(u_of(t)) for t in ts]
    w = [mp.mpf(1)] * nodes
    c = None
    for _ in range(passes):
        A = mp.matrix(nodes, deg + 1)
        b = mp.matrix(nodes, 1)
        for i, t in enumerate(ts):
            s = w[i] / f[i]
            base = t if times_t else mp.mpf(1)
            for j in range(deg + 1):
                A[i, j] = s * base * t ** j
            b[i] = s * f[i]
        c = mp.qr_solve(A, b)[0]
        err = []
        for i, t in enumerate(ts):
            p = sum(c[j] * t ** j for j in range(deg + 1)) * (t if times_t else 1)
            err.append(abs(p / f[i] - 1))
        # Lawson: weights grow where the error is large
        tot = sum(w[i] * err[i] for i in range(nodes))
        w = [w[i] * err[i] / tot * nodes for i in range(nodes)]
    return [c[j] for j in range(deg + 1)]


def horner(coef_low_first, t, dtype):
    """the kernels' Horner loop, each step one fused multiply-add rounded to dtype"""
    q = dtype(coef_low_first[-1])
    for c in reversed(coef_low_first[:-1]):
        q = dtype(np.float64(q) * np.float64(t) + np.float64(dtype(c)))
    return q


def max_rel_error(coef, times_t, u_hi, dtype, npts=20000):
    from scipy.special import erfcx
    u = np.concatenate([np.linspace(0.0, min(u_hi, 8.0), npts // 2),
                        np.geomspace(8.0, max(u_hi, 8.0 + 1e-9), npts // 2)]) if u_hi > 8 else \
        np.linspace(0.0, u_hi, npts)
    worst = 0.0
    for uu in u:
        t = dtype(1.0 / (1.0 + 0.5 * uu))
        p = horner(coef, t, dtype)
        if times_t:
            p = dtype(np.float64(t) * np.float64(p))
        worst = max(worst, abs(float(p) / erfcx(uu) - 1.0))
    return worst


def main():
    jobs = [("erfcx_tq", 10, True, t_of(mp.mpf(10000)), 1, np.float32, 1e4, 600, 12),
            ("erfcx_tq20", 20, True, mp.mpf("1e-8"), 1, np.float64, 1e4, 1200, 30)]
    for name, deg, times_t, lo, hi, dtype, u_hi, nodes, passes in jobs:
        c = fit(deg, times_t, lo, hi, nodes, passes)
        if dtype is np.float32:
            digits = ["%.9e" % float(np.float32(float(x))) for x in reversed(c)]
        else:
            digits = ["%.17e" % float(x) for x in reversed(c)]
        print(f"{name}: {'t * ' if times_t else ''}P{deg}(t), highest power first")
        print("  " + ", ".join(digits))
        print(f"  max relative error evaluated in {np.dtype(dtype).name} over u in [0, {u_hi:g}]: "
              f"{max_rel_error([float(x) for x in c], times_t, u_hi, dtype, 4000):.2e}")


if __name__ == "__main__":
    main()

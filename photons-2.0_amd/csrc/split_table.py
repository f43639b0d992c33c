"""Piecewise-polynomial tables of the split factor of the short-range pair force.

The reference multiplies each P2P pair force by G(u) = erfc(u) + 2u/sqrt(pi) exp(-u^2),
u = r / 2 rs (src/fmm.c:845-847, coeff = 2/sqrt(pi)).  The P2P kernels read G from a
table held in LDS instead of evaluating erfc and exp per pair: u in [0, U) is cut into
N equal intervals, and on interval i G is a polynomial in the fraction f in [0, 1) of
v = u N / U (i = floor(v), f = v - i):

    G(u) ~ c[i][0] + f (c[i][1] + f (c[i][2] + ...))

Each polynomial interpolates G at the Chebyshev nodes of its interval (computed in
40-digit arithmetic with mpmath, then rounded to the kernel's type); entry N is zero,
where the kernels clamp v.

    table   type    N    U    degree  |error| evaluated as the kernel does
    FP32    float   256  9.5  3       6e-8 (G <= 1); relative 5e-7 (u < 3), 6e-6 (u < 5)
    FP64    double  48   6    9       2e-15; relative 4e-15 (u < 3)

The FP32 table ends where exp(-u^2) leaves the float range (G(9.5) = 6e-39); the FP64
kernel evaluates G from erfc/exp beyond its table (G(6) = 1.6e-15, but the reference's
G is relatively exact wherever it is a normal double).

Run

    python photons-2.0_amd/csrc/split_table.py

to write split_table.h next to this script (a second or two);
tests/test_split_fits.py checks the header against this procedure and the error bounds
against mpmath's erfc.
"""
from __future__ import annotations

import os

import mpmath as mp

mp.mp.dps = 40

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "split_table.h")

# (name, C type, intervals, range of u, degree)
TABLES = [("kSplitTab32", "float", 256, 9.5, 3),
          ("kSplitTab64", "double", 48, 6, 9)]


def G(u):
    u = mp.mpf(u)
    return mp.erfc(u) + 2 * u / mp.sqrt(mp.pi) * mp.exp(-u * u)


def interval_poly(i, n, umax, deg):
    """monomial coefficients (lowest power first) in f of the polynomial interpolating
    G((i + f) umax / n) at the deg + 1 Chebyshev nodes of f in [0, 1]"""
    du = mp.mpf(umax) / n
    nodes = [(1 - mp.cos((2 * k + 1) * mp.pi / (2 * (deg + 1)))) / 2 for k in range(deg + 1)]
    V = mp.matrix([[x ** j for j in range(deg + 1)] for x in nodes])
    y = mp.matrix([G((i + x) * du) for x in nodes])
    c = mp.lu_solve(V, y)
    return [c[j] for j in range(deg + 1)]


def table(n, umax, deg):
    """n + 1 rows of deg + 1 coefficients (row n: zero)"""
    rows = [interval_poly(i, n, umax, deg) for i in range(n)]
    rows.append([mp.mpf(0)] * (deg + 1))
    return rows


def fmt(x, ctype):
    import numpy as np
    if ctype == "float":
        return "%.9ef" % float(np.float32(float(x)))
    return "%.17e" % float(x)


def render():
    out = ["// split_table.h -- written by split_table.py in this directory; do not edit.",
           "//",
           "// G(u) = erfc(u) + 2u/sqrt(pi) exp(-u^2), the split factor of the short-range pair",
           "// force, as N polynomials on u in [0, U): with v = min(u N / U, N), i = floor(v) and",
           "// f = v - i, G = c[i][0] + f (c[i][1] + f (c[i][2] + ...)); row N is zero.  Each",
           "// polynomial interpolates G at the Chebyshev nodes of its interval.",
           "#pragma once",
           "",
           "namespace photons {",
           ""]
    for name, ctype, n, umax, deg in TABLES:
        rows = table(n, umax, deg)
        out.append(f"constexpr int {name}N = {n};")
        out.append(f"constexpr double {name}U = {float(umax)!r};")
        out.append(f"constexpr int {name}Deg = {deg};")
        out.append(f"static __constant__ {ctype} {name}[{n + 1}][{deg + 1}] = {{")
        for r in rows:
            out.append("    {" + ", ".join(fmt(x, ctype) for x in r) + "},")
        out.append("};")
        out.append("")
    out.append("}  // namespace photons")
    return "\n".join(out) + "\n"


def main():
    with open(OUT, "w") as f:
        f.write(render())
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()

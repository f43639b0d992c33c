"""FP64 numpy restatement of the reference's PM long-range force (one rank).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product library
never uses it.

The reference (src/partmesh.c:800-1550 with src/conv.f90:171-229) on one rank is:
  1. CIC deposit with mesh points at cell centres: i = (int)(x N/L),
     w = (x - (i + 1/2) L/N) N/L, the neighbour on the side of w (i+1 when w > 0,
     else i-1 with w -> -w), weights (1-w, w) per axis; mesh += m * wx wy wz
     (src/partmesh.c:877-944), then mesh *= (N/L)^3 (:948-959); cells wrap
     periodically (the ghost exchange of :961-1240 on one rank);
  2. potential = IFFT(FFT(mesh) * gf), both transforms unnormalised, with
     gf = pref exp(-k^2 (2 pi rs / L)^2) / (sinc_x sinc_y sinc_z)^4 / k^2,
     pref = L^2 / (pi N^3), k the integer wave vector (components above N/2 taken
     as negative), sinc(a) = sin(pi a / N) / (pi a / N) (1 at a = 0), gf(0) = pref,
     and pi = 3.1415926 as truncated there (src/conv.f90:142, 171-229);
  3. per particle and axis, the 4th-order central difference
     f1 invx (p[+1] - p[-1]) - f2 invx (p[+2] - p[-2]), f1 = 4/3, f2 = 1/6,
     invx = N/(2L), at the 8 CIC points, interpolated with the CIC weights in the
     order of src/partmesh.c:1328-1550.  acc_pm has no G; it attracts.
"""
from __future__ import annotations

import numpy as np

PI_TRUNC = 3.1415926  # src/conv.f90:142


def cic_weights(pos, nside, boxsize):
    """cell index, neighbour index and weights per axis (src/partmesh.c:881-914)"""
    norm = nside / boxsize
    delta = 1.0 / norm
    i = (pos * norm).astype(np.int64)  # (int) truncation; pos >= 0
    w = (pos - (i + 0.5) * delta) * norm
    pos_side = w > 0
    ii = np.where(pos_side, i + 1, i - 1)
    w = np.where(pos_side, w, -w)
    return i, ii, w, 1.0 - w


def deposit(pos, mass, nside, boxsize):
    i, ii, w, wn = cic_weights(pos, nside, boxsize)
    mesh = np.zeros((nside, nside, nside))
    m = lambda a: a % nside
    corners = [((i, i, i), (wn, wn, wn)), ((ii, i, i), (w, wn, wn)), ((i, ii, i), (wn, w, wn)),
               ((i, i, ii), (wn, wn, w)), ((ii, ii, i), (w, w, wn)), ((ii, i, ii), (w, wn, w)),
               ((i, ii, ii), (wn, w, w)), ((ii, ii, ii), (w, w, w))]
    for (a, b, c), (wa, wb, wc) in corners:
        np.add.at(mesh, (m(a[:, 0]), m(b[:, 1]), m(c[:, 2])), mass * wa[:, 0] * wb[:, 1] * wc[:, 2])
    return mesh * (nside / boxsize) ** 3


def green(nside, boxsize, rs):
    """gf on the full (N, N, N) grid of integer wave vectors (src/conv.f90:171-219)"""
    half = nside // 2
    a = np.arange(nside)
    a = np.where(a > half, a - nside, a).astype(np.float64)

    def sinc(v):
        x = PI_TRUNC * v / nside
        return np.where(v == 0, 1.0, np.sin(x) / np.where(v == 0, 1.0, x))

    s = sinc(a)
    l, m, n = np.meshgrid(a, a, a, indexing="ij")
    k2 = l * l + m * m + n * n
    ismth2 = (2 * PI_TRUNC * rs / boxsize) ** 2
    pref = boxsize * boxsize / (PI_TRUNC * nside ** 3)
    ff = 1.0 / (s[:, None, None] * s[None, :, None] * s[None, None, :])
    with np.errstate(divide="ignore", invalid="ignore"):
        gf = pref * np.exp(-k2 * ismth2) * ff ** 4 / k2
    gf[0, 0, 0] = pref
    return gf


def potential(mesh, boxsize, rs):
    n = mesh.shape[0]
    spec = np.fft.fftn(mesh)  # unnormalised forward
    phi = np.fft.ifftn(spec * green(n, boxsize, rs)) * n ** 3  # unnormalised backward
    return phi.real


def gradient(phi, nside, boxsize):
    """4th-order central differences on the periodic mesh (src/partmesh.c:1254-1262, 1328-1550)"""
    invx = 0.5 * nside / boxsize
    f1, f2 = 4.0 / 3.0, 1.0 / 6.0
    g = []
    for ax in range(3):
        p1, m1 = np.roll(phi, -1, ax), np.roll(phi, 1, ax)
        p2, m2 = np.roll(phi, -2, ax), np.roll(phi, 2, ax)
        g.append(f1 * invx * (p1 - m1) - f2 * invx * (p2 - m2))
    return g


def interpolate(grad, pos, nside, boxsize):
    i, ii, w, wn = cic_weights(pos, nside, boxsize)
    m = lambda a: a % nside
    acc = np.zeros_like(pos)
    order = [((i, i, i), (wn, wn, wn)), ((ii, i, i), (w, wn, wn)), ((i, ii, i), (wn, w, wn)),
             ((ii, ii, i), (w, w, wn)), ((i, i, ii), (wn, wn, w)), ((ii, i, ii), (w, wn, w)),
             ((i, ii, ii), (wn, w, w)), ((ii, ii, ii), (w, w, w))]
    for ax in range(3):
        tot = np.zeros(len(pos))
        for (a, b, c), (wa, wb, wc) in order:
            tot = tot + wa[:, 0] * wb[:, 1] * wc[:, 2] * grad[ax][m(a[:, 0]), m(b[:, 1]), m(c[:, 2])]
        acc[:, ax] = tot
    return acc


def pm_force(pos, mass, nside, boxsize, rs):
    """acc_pm of every particle (no G), caller order"""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    phi = potential(deposit(pos, mass, nside, boxsize), boxsize, rs)
    return interpolate(gradient(phi, nside, boxsize), pos, nside, boxsize)


def long_range_point(r, mass, rs):
    """the long-range part of a point mass's attraction at distance r (no G): the
    complement of the short-range factor, m/r^2 [erf(r/2rs) - r/(sqrt(pi) rs) e^{-r^2/4rs^2}]"""
    from scipy.special import erf
    u = r / (2 * rs)
    return mass / r ** 2 * (erf(u) - 2 * u / np.sqrt(np.pi) * np.exp(-u * u))

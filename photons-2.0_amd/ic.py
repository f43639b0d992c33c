"""Seeded synthetic particle sets and force parameters for the short-range FMM path.

Units follow the reference: kpc/h for lengths, 1e10 Msun/h for masses, and
G = 43007.105732 (demo/lcdm_g2.run:28).  Random numbers are derived from the raw
64-bit output of numpy's PCG64 bit generator (whose stream is fixed), so the same
seed gives the same particles on every machine running this numpy.

Force parameters follow src/initial.c:316-345:
    splitRadius  rs  = 1.25 * L / NSIDE
    SoftenScale  eps = 0.03 * L / N**0.3333333
    cutoffRadius rc  = factor * rs   (the reference hard-codes 4.5)
    MASSPART         = Omega_m * 3 * 0.01 / (8 pi G) * L**3 / N   (src/initial.c:596)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, asdict

import numpy as np

GRAV_CONST = 43007.105732
OMEGA_M = 0.25


@dataclass
class ForceParams:
    boxsize: float
    mass: float
    soft: float
    rs: float
    rc: float
    theta: float = 0.4
    maxleaf: int = 32
    order: int = 3
    periodic: int = 1

    def as_dict(self):
        return asdict(self)


def force_params(n: int, boxsize: float, nside_pm: int, rc_factor: float = 6.0, theta: float = 0.4,
                 maxleaf: int = 32, order: int = 3, omega_m: float = OMEGA_M) -> ForceParams:
    rs = 1.25 * boxsize / float(nside_pm)
    soft = 0.03 * boxsize / math.pow(float(n), 0.3333333)
    mass = omega_m * 3 * 0.01 / (8.0 * math.pi * GRAV_CONST) * (boxsize ** 3) / float(n)
    return ForceParams(boxsize=boxsize, mass=mass, soft=soft, rs=rs, rc=rc_factor * rs,
                       theta=theta, maxleaf=maxleaf, order=order, periodic=1)


# --------------------------------------------------------------------------- random streams

def uniform(seed: int, n: int) -> np.ndarray:
    """n doubles in [0, 1) from the top 53 bits of PCG64 raw output."""
    if n == 0:
        return np.zeros(0)
    bits = np.random.PCG64(seed).random_raw(n)
    return (bits >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, n: int) -> np.ndarray:
    """n standard normals by the Box-Muller transform of uniform(seed, 2*ceil(n/2))."""
    m = (n + 1) // 2
    u = uniform(seed, 2 * m)
    u1 = 1.0 - u[:m]  # (0, 1]
    u2 = u[m:]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2.0 * np.pi * u2), r * np.sin(2.0 * np.pi * u2)])
    return z[:n]


def wrap(pos: np.ndarray, boxsize: float) -> np.ndarray:
    pos = np.mod(pos, boxsize)
    pos[pos >= boxsize] = 0.0  # np.mod can round up to exactly boxsize
    return pos


# --------------------------------------------------------------------------- particle sets

def lattice_gauss(nside: int, boxsize: float, sigma_cells: float = 0.2, seed: int = 0) -> np.ndarray:
    """Uniform lattice at (i + 1/2) * cell plus Gaussian displacements of sigma_cells cells."""
    cell = boxsize / nside
    g = (np.arange(nside, dtype=np.float64) + 0.5) * cell
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    pos += sigma_cells * cell * normal(seed, pos.size).reshape(-1, 3)
    return wrap(pos, boxsize)


def uniform_box(n: int, boxsize: float, seed: int) -> np.ndarray:
    return wrap(uniform(seed, 3 * n).reshape(n, 3) * boxsize, boxsize)


def sphere_counts(budget: int, seed: int, mmin: float = 32.0, mmax: float = 200000.0,
                  slope: float = 1.9) -> np.ndarray:
    """Member counts from dN/dM ~ M^-slope on [mmin, mmax], drawn until the budget is
    met; the last sphere is truncated to the remaining budget."""
    if budget <= 0:
        return np.zeros(0, dtype=np.int64)
    k = 1.0 - slope
    lo, hi = mmin ** k, mmax ** k
    counts = []
    total = 0
    block = 4096
    seed_off = 0
    while total < budget:
        u = uniform(seed * 1000003 + seed_off, block)
        seed_off += 1
        m = np.floor((lo + u * (hi - lo)) ** (1.0 / k)).astype(np.int64)
        m = np.maximum(m, int(mmin))
        for c in m:
            c = int(min(c, budget - total))
            counts.append(c)
            total += c
            if total >= budget:
                break
    return np.asarray(counts, dtype=np.int64)


def plummer_spheres(counts: np.ndarray, boxsize: float, seed_centres: int, seed_radii: int,
                    a0: float = 50.0, m0: float = 32.0, amax: float = 1000.0,
                    rtrunc: float = 10.0) -> np.ndarray:
    """Plummer spheres with scale radius a = a0 (M/m0)^(1/3) (clipped at amax), centred
    uniformly at random and truncated at rtrunc * a."""
    ns = len(counts)
    total = int(counts.sum())
    centres = uniform(seed_centres, 3 * ns).reshape(ns, 3) * boxsize
    a = np.minimum(a0 * np.cbrt(counts.astype(np.float64) / m0), amax)
    # cumulative Plummer mass fraction at the truncation radius
    fmax = rtrunc ** 3 / (rtrunc ** 2 + 1.0) ** 1.5
    u = uniform(seed_radii, 3 * total).reshape(3, total)
    uf = np.maximum(u[0] * fmax, 1e-300)
    r_over_a = 1.0 / np.sqrt(uf ** (-2.0 / 3.0) - 1.0)
    cost = 2.0 * u[1] - 1.0
    sint = np.sqrt(np.maximum(0.0, 1.0 - cost * cost))
    phi = 2.0 * np.pi * u[2]
    owner = np.repeat(np.arange(ns), counts)
    r = r_over_a * a[owner]
    pos = centres[owner] + np.stack([r * sint * np.cos(phi), r * sint * np.sin(phi), r * cost], axis=1)
    return wrap(pos, boxsize)


def clustered(n: int, boxsize: float, seed: int = 1, cluster_fraction: float = 0.5):
    """Half uniform, half in Plummer spheres (the clustered distribution of the
    256^3 workload, scaled to any n).  Returns (pos, nspheres)."""
    nc = int(round(n * cluster_fraction))
    nu = n - nc
    pos_u = uniform_box(nu, boxsize, seed)
    counts = sphere_counts(nc, seed + 1)
    pos_c = plummer_spheres(counts, boxsize, seed + 2, seed + 3)
    pos = np.concatenate([pos_u, pos_c], axis=0)
    return pos, len(counts)


# --------------------------------------------------------------------------- named workloads

def workload_lattice32(rc_factor: float = 6.0, order: int = 3):
    """32^3 particles, 100 Mpc/h box, lattice + 0.2-cell Gaussian displacements,
    PM grid 64^3, theta 0.4, leaf 32."""
    n, box = 32 ** 3, 100000.0
    pos = lattice_gauss(32, box, 0.2, seed=0)
    return pos, force_params(n, box, 64, rc_factor=rc_factor, order=order)


def workload_clustered(nside: int = 256, box_at_256: float = 500000.0, rc_factor: float = 6.0,
                       order: int = 3, theta: float = 0.4, seed: int = 1):
    """nside^3 clustered particles; the box scales with N so that the mean density is
    the one of 256^3 particles in 500 Mpc/h; PM grid nside^3."""
    n = nside ** 3
    box = box_at_256 * (n / 256.0 ** 3) ** (1.0 / 3.0)
    pos, _ = clustered(n, box, seed=seed)
    return pos, force_params(n, box, nside, rc_factor=rc_factor, order=order, theta=theta)


def ensemble_sizes(nbox: int, nmin: int, nmax: int, seed: int) -> np.ndarray:
    u = uniform(seed, nbox)
    return np.floor(np.exp(np.log(nmin) + u * (np.log(nmax) - np.log(nmin)))).astype(np.int64)


def workload_ensemble(nbox: int = 64, nmin: int = 32 ** 3, nmax: int = 128 ** 3, seed: int = 5,
                      order: int = 4, theta: float = 0.4, rc_factor: float = 6.0, box_at_256: float = 500000.0):
    """Independent periodic boxes of log-uniform sizes in [nmin, nmax] (seeded), each
    clustered like the 256^3 workload at its density: L_b = 500 Mpc/h (N_b/256^3)^(1/3).
    At equal density rs (1.25 mean spacings), the softening and the particle mass are
    the same for every box, so one parameter set serves the batch.
    Returns (counts, lengths, concatenated positions, shared ForceParams)."""
    counts = ensemble_sizes(nbox, nmin, nmax, seed)
    lengths = box_at_256 * (counts / 256.0 ** 3) ** (1.0 / 3.0)
    pos = [clustered(int(n), float(L), seed=1000 + 10 * b)[0] for b, (n, L) in enumerate(zip(counts, lengths))]
    prm = force_params(256 ** 3, box_at_256, 256, rc_factor=rc_factor, order=order, theta=theta)
    return counts, lengths, np.concatenate(pos, axis=0), prm

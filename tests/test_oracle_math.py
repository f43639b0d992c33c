"""Mathematical pins of the CPU oracle that do not depend on the reference.

They cover what the reference cannot (orders 4 and 5, whose reference M2L is
incomplete) and would catch a dropped term, a wrong sign or a transposed index:
  - radial factors against sympy derivatives of erfc(r/2rs)/r,
  - derivative tensors against sympy partial derivatives,
  - translation identities (M2M o P2M, L2P o L2L),
  - L2P against the analytic gradient of the local expansion,
  - M2L + L2P against the direct pair sum, converging with the order,
  - the pair kernel against -grad of the split potential and its limits,
  - acceptance() on boxes whose answer is fixed by hand,
  - tree invariants and the exact-once coverage of every pair inside rc,
  - the whole force call against a brute-force direct sum.
"""
import math

import numpy as np
import pytest
import sympy as sp

import ic
import oracle
from cases import case_inputs


def multi_indices(order):
    out = []
    for n in range(order + 1):
        for a in range(n, -1, -1):
            for b in range(n - a, -1, -1):
                out.append((a, b, n - a - b))
    return out


def test_index_order_is_the_reference_layout():
    mi = multi_indices(4)
    # inc/operator.h: X=1, XX=4, XY=5, YY=7, ZZ=9, XXX=10, XYZ=14, ZZZ=19, XXXX=20, XXYY=23, YYYY=30, ZZZZ=34
    for name, k in (("X", 1), ("XX", 4), ("XY", 5), ("YY", 7), ("ZZ", 9), ("XXX", 10), ("XYZ", 14),
                    ("ZZZ", 19), ("XXXX", 20), ("XXYY", 23), ("YYYY", 30), ("ZZZZ", 34)):
        assert mi[k] == (name.count("X"), name.count("Y"), name.count("Z"))
    assert [oracle.nterms(p) for p in (3, 4, 5)] == [20, 35, 56]


# --------------------------------------------------------------------- kernels

R_SYM, RS_SYM = sp.symbols("r rs", positive=True)
G_SYM = sp.erfc(R_SYM / (2 * RS_SYM)) / R_SYM


@pytest.mark.parametrize("r_over_rs", [0.05, 0.7, 2.0, 4.5, 6.0])
def test_radial_factors_against_symbolic_derivatives(r_over_rs):
    rs = 1953.125
    r = r_over_rs * rs
    fac = oracle.radial(r, rs, 6)
    f = G_SYM
    for n in range(7):
        exact = float(f.subs({R_SYM: r, RS_SYM: rs}).evalf(40))
        assert abs(fac[n] - exact) <= 1e-12 * abs(exact), (n, fac[n], exact)
        f = sp.diff(f, R_SYM) / R_SYM


@pytest.mark.parametrize("order", [3, 4, 5])
def test_derivative_tensor_against_symbolic_partials(order):
    x, y, z = sp.symbols("x y z", real=True)
    rs = 1500.0
    g = sp.erfc(sp.sqrt(x * x + y * y + z * z) / (2 * rs)) / sp.sqrt(x * x + y * y + z * z)
    R = np.array([1234.5, -2222.0, 987.25])
    D = oracle.dtensor(order, rs, R)
    for k, (a, b, c) in enumerate(multi_indices(order)):
        expr = g
        for v, m in ((x, a), (y, b), (z, c)):
            if m:
                expr = sp.diff(expr, v, m)
        exact = float(expr.subs({x: R[0], y: R[1], z: R[2]}).evalf(40))
        scale = abs(float(D[0])) / np.linalg.norm(R) ** (a + b + c)
        assert abs(D[k] - exact) <= 1e-12 * max(abs(exact), scale), ((a, b, c), D[k], exact)


def test_pair_kernel_limits_and_gradient():
    m, eps, rs = 211.7, 93.75, 1953.125
    xt = np.array([1000.0, 2000.0, 3000.0])
    # softened branch, r < eps: k = m/eps^3 * G(u)
    d = np.array([20.0, -30.0, 10.0])
    r = np.linalg.norm(d)
    u = 0.5 * r / rs
    G = math.erfc(u) + 2.0 / math.sqrt(math.pi) * u * math.exp(-u * u)
    np.testing.assert_allclose(oracle.p2p_pair(xt, xt + d, m, eps, rs), d * m / eps ** 3 * G, rtol=1e-14)
    # outside the softening: the force is grad_target of m*erfc(r/2rs)/r, by central differences
    d = np.array([700.0, -1200.0, 450.0])
    a = oracle.p2p_pair(xt, xt + d, m, eps, rs)
    phi = lambda p: m * math.erfc(np.linalg.norm(xt + d - p) / (2 * rs)) / np.linalg.norm(xt + d - p)
    h = 1e-2
    num = np.array([(phi(xt + h * e) - phi(xt - h * e)) / (2 * h) for e in np.eye(3)])
    np.testing.assert_allclose(a, num, rtol=1e-7)
    # Newtonian limit r << rs and the erfc tail far beyond rs
    d = np.array([1e-2 * rs, 0.0, 0.0])
    np.testing.assert_allclose(oracle.p2p_pair(xt, xt + d, m, 1e-6, rs), m * d / np.linalg.norm(d) ** 3, rtol=1e-4)
    d = np.array([12.0 * rs, 0.0, 0.0])
    assert np.linalg.norm(oracle.p2p_pair(xt, xt + d, m, eps, rs)) < 1e-14 * m / (12.0 * rs) ** 2


# --------------------------------------------------------------------- translations

@pytest.mark.parametrize("order", [3, 4, 5])
def test_m2m_of_p2m_is_p2m_at_parent(order):
    rng = np.random.Generator(np.random.PCG64(order))
    pos = rng.uniform(-400.0, 400.0, (25, 3)) + np.array([5000.0, 6000.0, 7000.0])
    child = np.array([5100.0, 5950.0, 7030.0])
    parent = np.array([5400.0, 5800.0, 7500.0])
    Mc = oracle.p2m(order, pos, 2.5, child)
    Mp = oracle.m2m(order, parent - child, Mc)
    np.testing.assert_allclose(Mp, oracle.p2m(order, pos, 2.5, parent), rtol=1e-11,
                               atol=1e-13 * np.abs(Mp).max())


@pytest.mark.parametrize("order", [3, 4, 5])
def test_l2l_is_exact_for_the_local_polynomial(order):
    rng = np.random.Generator(np.random.PCG64(10 + order))
    nt = oracle.nterms(order)
    deg = np.array([sum(m) for m in multi_indices(order)])
    L = rng.normal(size=nt) * 10.0 ** (-3.0 * deg)
    parent = np.array([0.0, 0.0, 0.0])
    child = np.array([300.0, -200.0, 150.0])
    Lc = oracle.l2l(order, child - parent, L)
    pts = child + rng.uniform(-200.0, 200.0, (16, 3))
    np.testing.assert_allclose(oracle.l2p(order, Lc, child, pts), oracle.l2p(order, L, parent, pts),
                               rtol=1e-11, atol=1e-14 * np.abs(L[1:4]).max())


@pytest.mark.parametrize("order", [3, 4, 5])
def test_l2p_is_the_gradient_of_the_local_expansion(order):
    rng = np.random.Generator(np.random.PCG64(20 + order))
    mi = multi_indices(order)
    L = rng.normal(size=len(mi))
    c = np.array([10.0, 20.0, 30.0])
    pts = c + rng.uniform(-1.0, 1.0, (8, 3))
    acc = oracle.l2p(order, L, c, pts)
    for p, a in zip(pts, acc):
        d = p - c
        grad = np.zeros(3)
        for Lk, (i, j, k) in zip(L, mi):
            f = math.factorial(i) * math.factorial(j) * math.factorial(k)
            if i: grad[0] += Lk * i * d[0] ** (i - 1) * d[1] ** j * d[2] ** k / f
            if j: grad[1] += Lk * j * d[0] ** i * d[1] ** (j - 1) * d[2] ** k / f
            if k: grad[2] += Lk * k * d[0] ** i * d[1] ** j * d[2] ** (k - 1) / f
        np.testing.assert_allclose(a, grad, rtol=1e-12, atol=1e-12)


def test_m2l_then_l2p_converges_to_the_direct_sum():
    rng = np.random.Generator(np.random.PCG64(77))
    rs, m = 2000.0, 3.0
    src_c = np.array([0.0, 0.0, 0.0])
    tgt_c = np.array([4200.0, -2500.0, 1800.0])
    src = src_c + rng.uniform(-300.0, 300.0, (30, 3))
    tgt = tgt_c + rng.uniform(-300.0, 300.0, (10, 3))
    direct = np.zeros((len(tgt), 3))
    for i, xt in enumerate(tgt):
        for xs in src:
            direct[i] += oracle.p2p_pair(xt, xs, m, 1e-3, rs)
    errs = []
    for order in (1, 2, 3, 4, 5):
        M = oracle.p2m(order, src, m, src_c)
        L = oracle.m2l(order, rs, tgt_c - src_c, M)
        acc = oracle.l2p(order, L, tgt_c, tgt)
        errs.append(np.abs(acc - direct).max() / np.abs(direct).max())
    # (source + target radius) / distance is about 0.2: ~10x per order
    assert errs[-1] < 5e-5, errs
    assert all(b < 0.2 * a for a, b in zip(errs, errs[1:])), errs


# --------------------------------------------------------------------- acceptance

@pytest.mark.parametrize("wi,wj,dist,theta,rc,expect", [
    ((100, 100, 100), (100, 100, 100), (100, 0, 0), 0.4, 1e5, 0),          # face-touching boxes
    ((100, 100, 100), (100, 100, 100), (100 + 5e-5, 0, 0), 0.4, 1e5, 0),   # gap below 1e-4 kpc/h
    ((100, 100, 100), (100, 100, 100), (100 + 2e-4, 0, 0), 0.4, 1e5, 0),   # separated, MAC fails
    ((10, 10, 10), (10, 10, 10), (1000, 0, 0), 0.4, 500, -1),               # gap beyond the cutoff
    ((10, 10, 10), (10, 10, 10), (1000, 0, 0), 0.4, 5000, 1),               # small and far: accept
    ((200, 200, 200), (200, 200, 200), (1100, 0, 0), 0.4, 1050, 0),         # centres beyond rc, gap within
    ((50, 10, 10), (50, 10, 10), (0, 300, 0), 0.4, 5000, 1),                # (2*50)^2 < 0.16*300^2
    ((61, 10, 10), (61, 10, 10), (0, 300, 0), 0.4, 5000, 0),                # (2*61)^2 > 0.16*300^2
    ((10, 10, 10), (10, 10, 10), (-800, 600, 0), 0.4, 5000, 1),             # sign of dist is irrelevant
])
def test_acceptance_truth_table(wi, wj, dist, theta, rc, expect):
    assert oracle.acceptance(wi, wj, dist, theta, rc) == expect


# --------------------------------------------------------------------- tree and walk

def _cell_ranges(ex, nleaf):
    """particle range [lo, hi) of every cell (leaves first, then nodes)"""
    lo = np.concatenate([ex["leaf_ipart"], np.zeros(len(ex["node_npart"]), np.int64)]).astype(np.int64)
    hi = np.concatenate([ex["leaf_ipart"] + ex["leaf_npart"], np.zeros(len(ex["node_npart"]), np.int64)])
    son = ex["node_son"]
    for k in range(len(son) - 1, -1, -1):  # children have larger pre-order ids
        cl = [(-s - 1) if s < 0 else nleaf + s for s in son[k]]
        lo[nleaf + k] = min(lo[c] for c in cl)
        hi[nleaf + k] = max(hi[c] for c in cl)
    return lo, hi


def test_tree_invariants():
    pos, prm = case_inputs("clus24_pm48")
    t = oracle.Tree(pos, prm)
    ex = t.export()
    n = len(pos)
    assert sorted(ex["perm"]) == list(range(n))
    p = pos[ex["perm"]]
    ip, np_ = ex["leaf_ipart"], ex["leaf_npart"]
    assert ip[0] == 0 and np.all(ip[1:] == ip[:-1] + np_[:-1]) and ip[-1] + np_[-1] == n
    assert np_.max() <= prm.maxleaf
    # every particle lies in its leaf's box
    for l in range(t.nleaf):
        q = p[ip[l]: ip[l] + np_[l]]
        lo = ex["leaf_center"][l] - 0.5 * ex["leaf_width"][l]
        hi = ex["leaf_center"][l] + 0.5 * ex["leaf_width"][l]
        assert np.all(q >= lo - 1e-9) and np.all(q <= hi + 1e-9)
    # split = mean of the node's particles along its axis; left <= split < right
    lo, hi = _cell_ranges(ex, t.nleaf)
    depth = np.zeros(t.nnode, np.int64)
    for k in range(t.nnode):
        axis = depth[k] % 3
        seg = p[lo[t.nleaf + k]: hi[t.nleaf + k], axis]
        assert abs(ex["node_split"][k] - seg.mean()) <= 1e-9 * prm.boxsize
        s0, s1 = ex["node_son"][k]
        c0 = (-s0 - 1) if s0 < 0 else t.nleaf + s0
        assert np.all(p[lo[c0]: hi[c0], axis] <= ex["node_split"][k])
        c1 = (-s1 - 1) if s1 < 0 else t.nleaf + s1
        assert np.all(p[lo[c1]: hi[c1], axis] > ex["node_split"][k])
        for s in (s0, s1):
            if s >= 0:
                assert s > k
                depth[s] = depth[k] + 1


SHIFTS = np.array([(i, j, k) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)
                   if (i, j, k) != (0, 0, 0)], np.float64)


def _expand(lo, hi, a, b, s, n):
    """all (i, j, shift) keys of the cell pairs (a, b, s), vectorised"""
    nt = hi[a] - lo[a]
    ns = hi[b] - lo[b]
    cnt = nt * ns
    e = np.repeat(np.arange(len(a)), cnt)
    off = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    i = lo[a][e] + off // ns[e]
    j = lo[b][e] + off % ns[e]
    sh = s[e].astype(np.int64)
    keep = ~((i == j) & (sh == 0))
    return ((i * n + j) * 27 + sh)[keep]


@pytest.mark.parametrize("name,theta", [("clus16_pm32", 0.4), ("clus16_pm32", 0.3), ("clus16_pm32", 0.6)])
def test_every_pair_inside_the_cutoff_is_covered_exactly_once(name, theta):
    pos, prm = case_inputs(name)
    prm.theta = theta
    t = oracle.Tree(pos, prm)
    ex = t.export()
    n = len(pos)
    p = pos[ex["perm"]]
    lo, hi = _cell_ranges(ex, t.nleaf)
    pt, ps, psh = t.p2p_list()
    mt, ms, msh = t.m2l_list()
    assert len(mt) > 0
    keys = np.concatenate([_expand(lo, hi, pt, ps, psh, n), _expand(lo, hi, mt, ms, msh, n)])
    uniq, cnt = np.unique(keys, return_counts=True)
    assert cnt.max() == 1, "a pair is interacted twice"
    shifts = np.vstack([np.zeros(3), SHIFTS]) * prm.boxsize
    need = []
    for s in range(27):
        d = p[None, :, :] + shifts[s][None, None, :] - p[:, None, :]
        r2 = (d * d).sum(axis=2)
        I, J = np.nonzero(r2 < prm.rc * prm.rc)
        sel = (I != J) if s == 0 else np.ones(len(I), bool)
        need.append((I[sel].astype(np.int64) * n + J[sel]) * 27 + s)
    need = np.concatenate(need)
    assert np.isin(need, uniq).all(), "a pair inside the cutoff is never interacted"


def test_force_call_against_direct_sum():
    n = 16 ** 3
    box = 500000.0 * (n / 256.0 ** 3) ** (1.0 / 3.0)
    pos, _ = ic.clustered(n, box, seed=1)
    prm = ic.force_params(n, box, 32, rc_factor=8.0, order=5, theta=0.2)
    acc, st = oracle.force(pos, prm)
    assert st["m2l_calls"] > 0
    ref = oracle.direct(pos, prm, np.arange(n), prm.rc)
    err = np.linalg.norm(acc - ref, axis=1)
    mag = np.linalg.norm(ref, axis=1)
    assert np.percentile(err / mag, 99) < 1e-5
    assert err.max() < 1e-3 * np.median(mag)


def test_force_on_selected_leaves_is_the_whole_force_there():
    """orc_force_leaves (the headline-size parity checker) against orc_force: the same
    walks entered only above the selected leaves give those leaves' particles the same
    force bit for bit, and the same list lengths as the full lists"""
    from cases import case_inputs
    pos, prm = case_inputs("clus24_pm48")
    prm.order = 4
    full, _ = oracle.force(pos, prm)
    t = oracle.Tree(pos, prm)
    sel = np.random.default_rng(4).choice(t.nleaf, 25, replace=False).astype(np.int32)
    acc, ids, plen, mlen = oracle.force_leaves(pos, prm, sel)
    ex = t.export()
    assert len(ids) == ex["leaf_npart"][sel].sum()
    np.testing.assert_array_equal(acc, full[ids])
    tg, _, _ = t.p2p_list()
    np.testing.assert_array_equal(plen, np.bincount(tg, minlength=t.nleaf)[sel])
    # M2L entries of each leaf and its ancestors, from the full list
    mt, _, _ = t.m2l_list()
    per_cell = np.bincount(mt, minlength=t.nleaf + t.nnode)
    parent = np.full(t.nleaf + t.nnode, -1)
    for v, (a, b) in enumerate(ex["node_son"]):
        for c in (a, b):
            parent[c + t.nleaf if c >= 0 else -c - 1] = t.nleaf + v
    for k, leaf in enumerate(sel):
        tot, c = 0, int(leaf)
        while c >= 0:
            tot += per_cell[c]
            c = parent[c]
        assert mlen[k] == tot

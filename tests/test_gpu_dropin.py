"""The drop-in layer: the reference's globals and entry points, driven like the reference.

The reference driver calls fmm_construct, fmm_prepare, fmm_task and fmm_ext on the
global particle array `part` (src/photoNs.c:97-116) and reads part[].acc; the force
call permutes part[] (src/fmm.c:93), so particles are matched through a tag kept in
vel[0], as the reference harness does.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
import photons_gpu as pg
from cases import case_inputs

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


class Ref:
    """the reference globals of the library"""

    def __init__(self):
        self.lib = pg.lib()
        for f in ("fmm_construct", "fmm_prepare", "fmm_task", "fmm_ext", "fmm_deconstruct"):
            getattr(self.lib, f).restype = None
        dp = C.POINTER(C.c_double)
        self.lib.p2m.argtypes = [C.c_int, C.c_int, dp, dp]
        self.lib.m2m.argtypes = [C.c_double, C.c_double, C.c_double, dp, dp]
        self.lib.m2l.argtypes = [C.c_double, C.c_double, C.c_double, dp, dp]
        self.lib.l2l.argtypes = [C.c_double, C.c_double, C.c_double, dp, dp]
        self.lib.acceptance.argtypes = [dp, dp, dp]
        self.lib.fmm_remote.argtypes = [C.c_int, dp]
        for f in ("p2p_kernel", "p2p_kernel_inleaf", "p2p_kernel_rinvsqrt", "walk_task_p2p", "walk_task_m2l",
                  "walk_task_p2p_ext", "walk_task_m2l_ext"):
            getattr(self.lib, f).argtypes = [C.c_int, C.c_int]
        self.lib.l2p.argtypes = [C.c_int]
        self.lib.walk_l2l.argtypes = [C.c_int]
        self.lib.walk_m2m.argtypes = [C.c_int]

    def i(self, name):
        return C.c_int.in_dll(self.lib, name)

    def d(self, name):
        return C.c_double.in_dll(self.lib, name)

    def u(self, name):
        return C.c_ulong.in_dll(self.lib, name)

    def set_params(self, prm, n):
        self.i("NPART").value = n
        self.i("MAXLEAF").value = prm.maxleaf
        self.d("MASSPART").value = prm.mass
        self.d("SoftenScale").value = prm.soft
        self.d("splitRadius").value = prm.rs
        self.d("cutoffRadius").value = prm.rc
        self.d("open_angle").value = prm.theta
        self.d("BOXSIZE").value = prm.boxsize
        # counters the reference never resets itself (the harness that wrote the
        # golden vectors starts them at zero)
        self.u("walk_m2l_count").value = 0
        self.u("p2p_count_remote").value = 0

    def set_part(self, body):
        self.body = body  # keep alive
        C.c_void_p.in_dll(self.lib, "part").value = body.ctypes.data


def bodies(pos):
    b = np.zeros((len(pos), 12))
    b[:, 0:3] = pos
    b[:, 6] = np.arange(len(pos))
    return b


def unpermute(body):
    idx = body[:, 6].astype(np.int64)
    acc = np.zeros((len(body), 3))
    acc[idx] = body[:, 3:6]
    return acc


def check_golden(name, acc):
    g = np.load(os.path.join(GOLDEN, f"force_{name}.npz"), allow_pickle=False)
    e = np.linalg.norm(acc[g["idx"]] - g["acc"], axis=1) / np.linalg.norm(g["acc"], axis=1)
    assert e.max() <= 1e-10, e.max()
    return g


@pytest.mark.parametrize("name", ["clus24_pm48", "lattice32_rc4.5"])
def test_force_call_as_the_driver_makes_it(name):
    r = Ref()
    pos, prm = case_inputs(name)
    r.set_params(prm, len(pos))
    body = bodies(pos)
    r.set_part(body)
    r.lib.fmm_construct()
    r.lib.fmm_prepare()
    # part[] now in tree order: leaf particle ranges are contiguous
    perm = body[:, 6].astype(np.int64)
    assert sorted(perm) == list(range(len(pos)))
    np.testing.assert_array_equal(body[:, 0:3], pos[perm])
    r.lib.fmm_task()
    r.lib.fmm_ext()
    g = check_golden(name, unpermute(body))
    c = g["counters"]
    assert r.u("idxP2P").value == c[0] and r.u("idxM2L").value == c[1]
    assert r.u("walk_m2l_count").value == c[2] and r.u("p2p_count_remote").value == c[3]
    assert r.i("last_leaf").value - r.i("first_leaf").value == c[4]
    assert r.i("last_node").value - r.i("first_node").value + 1 == c[5]
    r.lib.fmm_deconstruct()


def test_reference_sequence_through_the_single_entry_points():
    """fmm_task and fmm_ext restated with the reference's own inner calls
    (src/fmm.c:772-793, 909-984, 1028-1057)."""
    r = Ref()
    name = "clus16_pm32"
    pos, prm = case_inputs(name)
    r.set_params(prm, len(pos))
    body = bodies(pos)
    r.set_part(body)
    r.lib.fmm_construct()
    r.lib.fmm_prepare()
    root = r.i("first_node").value
    r.lib.walk_task_p2p(root, root)
    r.lib.walk_task_m2l(root, root)
    L = prm.boxsize
    for mi in (-1, 0, 1):
        for mj in (-1, 0, 1):
            for mk in (-1, 0, 1):
                if mi == mj == mk == 0:
                    continue
                sh = (C.c_double * 3)(mi * L, mj * L, mk * L)
                r.lib.fmm_remote(0, sh)
    r.lib.walk_l2l(root)
    for leaf in range(r.i("first_leaf").value, r.i("last_leaf").value):
        r.lib.l2p(leaf)
    check_golden(name, unpermute(body))


def test_single_operators_match_reference_golden():
    r = Ref()
    _, prm = case_inputs("clus16_pm32")
    r.set_params(prm, 0)
    g = np.load(os.path.join(GOLDEN, "ops.npz"), allow_pickle=False)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    # p2m over part[]
    pos = np.split(g["p2m_pos"], np.cumsum(g["p2m_n"])[:-1])
    for k, p in enumerate(pos):
        body = bodies(p)
        r.set_part(body)
        r.d("MASSPART").value = g["p2m_mass"][k]
        M = np.zeros(20)
        c = np.ascontiguousarray(g["p2m_center"][k])
        r.lib.p2m(0, len(p), dp(c), dp(M))
        np.testing.assert_allclose(M, g["p2m_M"][k], rtol=1e-12, atol=1e-14 * np.abs(g["p2m_M"][k]).max())
    for op in ("m2m", "l2l", "m2l"):
        for k in range(len(g[f"{op}_d"])):
            d = g[f"{op}_d"][k]
            X = np.ascontiguousarray(g[f"{op}_in"][k])
            Y = np.zeros(20)
            if op == "m2l":
                r.d("splitRadius").value = g["m2l_rs"][k]
            getattr(r.lib, op)(d[0], d[1], d[2], dp(X), dp(Y))
            ref = g[f"{op}_out"][k]
            for lo, hi in ((0, 1), (1, 4), (4, 10), (10, 20)):
                assert np.abs(Y[lo:hi] - ref[lo:hi]).max() <= 1e-12 * np.abs(ref[lo:hi]).max(), (op, k)
    for row, flag in zip(g["acc_in"], g["acc_flag"]):
        r.d("open_angle").value = row[0]
        r.d("cutoffRadius").value = row[1]
        wi, wj, di = (np.ascontiguousarray(row[a:a + 3]) for a in (2, 5, 8))
        assert r.lib.acceptance(dp(wi), dp(wj), dp(di)) == flag


class RefNode(C.Structure):
    """Node of inc/typesdef.h:25-36 (p = 3), as the library mirrors btree[]"""
    _fields_ = [("updated", C.c_int), ("npart", C.c_int), ("son", C.c_int * 2), ("split", C.c_double),
                ("width", C.c_double * 3), ("center", C.c_double * 3), ("M", C.c_double * 20),
                ("L", C.c_double * 20)]


def test_walk_m2m_matches_reference_multipoles():
    """walk_m2m(first_node) (src/operator.c:165-194) through the drop-in layer on node
    mirrors whose multipoles were zeroed: every node's M equals the one the reference's
    own fmm_prepare computed (tests/golden/tree_clus16_pm32.npz), and so does the M the
    GPU's M2M pass left there before"""
    g = np.load(os.path.join(GOLDEN, "tree_clus16_pm32.npz"), allow_pickle=False)
    r = Ref()
    pos, prm = case_inputs("clus16_pm32")
    r.set_params(prm, len(pos))
    body = bodies(pos)
    r.set_part(body)
    r.lib.fmm_construct()
    r.lib.fmm_prepare()
    fn, ln = r.i("first_node").value, r.i("last_node").value
    nnode = ln - fn + 1
    assert nnode == len(g["node_M"])
    base = C.c_void_p.in_dll(r.lib, "btree").value + fn * C.sizeof(RefNode)
    nodes = (RefNode * nnode).from_address(base)
    ref = g["node_M"]

    def check(M):
        for lo, hi in ((0, 1), (1, 4), (4, 10), (10, 20)):
            scale = np.abs(ref[:, lo:hi]).max()
            assert np.abs(M[:, lo:hi] - ref[:, lo:hi]).max() <= 1e-12 * scale, (lo, hi)

    check(np.array([list(nd.M) for nd in nodes]))
    for nd in nodes:
        for q in range(20):
            nd.M[q] = 0.0
    r.lib.walk_m2m(fn)
    check(np.array([list(nd.M) for nd in nodes]))
    r.lib.fmm_deconstruct()


def test_pair_kernels_of_kernels_h():
    r = Ref()
    pos, prm = case_inputs("clus16_pm32")
    r.set_params(prm, len(pos))
    body = bodies(pos)
    r.set_part(body)
    r.lib.fmm_construct()
    r.lib.fmm_prepare()
    fl = r.i("first_leaf").value
    t = oracle.Tree(pos, prm)
    ex = t.export()
    # a populated pair of distinct leaves and one leaf with itself
    big = np.argsort(-ex["leaf_npart"])[:2]
    i, j = int(big[0]), int(big[1])
    tree_pos = body[:, 0:3].copy()
    for kern, src in (("p2p_kernel", j), ("p2p_kernel_inleaf", i), ("p2p_kernel_rinvsqrt", j)):
        body[:, 3:6] = 0.0
        getattr(r.lib, kern)(fl + i, fl + src)
        a, b = ex["leaf_ipart"][i], ex["leaf_ipart"][i] + ex["leaf_npart"][i]
        sa, sb = ex["leaf_ipart"][src], ex["leaf_ipart"][src] + ex["leaf_npart"][src]
        want = np.zeros((b - a, 3))
        for q in range(a, b):
            for s in range(sa, sb):
                if s != q:
                    want[q - a] += oracle.p2p_pair(tree_pos[q], tree_pos[s], prm.mass, prm.soft, prm.rs)
        got = body[a:b, 3:6]
        tol = 1e-5 if kern.endswith("rinvsqrt") else 1e-12
        assert np.abs(got - want).max() <= tol * np.abs(want).max(), kern
        assert np.abs(body[:a, 3:6]).max(initial=0.0) == 0.0


DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "photons-2.0_amd", "examples",
                      "fmm_driver")


def run_driver(pos, prm, tmp, fp32=False):
    """the C program photons-2.0_amd/examples/fmm_driver.c: the reference driver's
    force call (src/photoNs.c:222-252) linked against libphotons_fmm.so"""
    import subprocess
    fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
    d = prm.as_dict()
    with open(fin, "wb") as f:
        f.write(np.array([len(pos), d["maxleaf"], 0, 0], np.int32).tobytes())
        f.write(np.array([d["boxsize"], d["mass"], d["soft"], d["rs"], d["rc"], d["theta"]]).tobytes())
        f.write(np.ascontiguousarray(pos, dtype=np.float64).tobytes())
    subprocess.run([DRIVER, fin, fout] + (["fp32"] if fp32 else []), check=True, timeout=120)
    raw = open(fout, "rb").read()
    n = len(pos)
    acc = np.frombuffer(raw[: 24 * n], dtype=np.float64).reshape(n, 3)
    cnt = np.frombuffer(raw[24 * n: 24 * n + 48], dtype=np.int64)
    tail = np.frombuffer(raw[24 * n + 48:], dtype=np.float64)
    timers = dict(zip(("prep", "task", "ext", "fmm", "p2p", "m2l"), tail[:6]))
    return acc, cnt, timers, tail[6:]


@pytest.mark.parametrize("name", ["clus16_pm32", "lattice32_rc6"])
def test_compiled_c_driver_matches_reference_golden(name, tmp_path):
    pos, prm = case_inputs(name)
    acc, cnt, _, _ = run_driver(pos, prm, str(tmp_path))
    g = check_golden(name, acc)
    assert list(cnt[:4]) == list(g["counters"][:4])
    assert cnt[4] == g["counters"][4] and cnt[5] == g["counters"][5]
    a32, _, _, _ = run_driver(pos, prm, str(tmp_path), fp32=True)
    e = np.linalg.norm(a32[g["idx"]] - g["acc"], axis=1) / np.linalg.norm(g["acc"], axis=1)
    assert np.percentile(e, 99) <= 1e-4


def test_partmesh_drop_in_matches_the_pm_oracle():
    """partmesh() with the reference's globals (src/partmesh.c:800-1550) fills
    part[].acc_pm; checked against the numpy restatement oracle/pm.py"""
    import pm
    r = Ref()
    pos, prm = case_inputs("lattice32_rc6")
    r.set_params(prm, len(pos))
    r.i("NSIDE").value = 64
    body = bodies(pos)
    r.set_part(body)
    r.lib.partmesh.restype = None
    r.lib.partmesh()
    want = pm.pm_force(pos, prm.mass, 64, prm.boxsize, prm.rs)
    got = body[:, 9:12]
    assert np.abs(got - want).max() <= 1e-11 * np.abs(want).max()
    # the FMM globals and part[].pos are untouched
    np.testing.assert_array_equal(body[:, 0:3], pos)


def test_reference_timers_are_written(tmp_path):
    """dtime_prep/task/ext/fmm/p2p/m2l (inc/photoNs.h:95-111) as the reference driver
    prints them (src/photoNs.c:122-127): non-zero, the kernel timers equal the library's
    phase times of the same call, and the phases of fmm_task fit inside its wall time"""
    pos, prm = case_inputs("clus24_pm48")
    _, _, t, ph = run_driver(pos, prm, str(tmp_path))
    print(t, ph)
    for k in ("prep", "task", "ext", "fmm", "p2p", "m2l"):
        assert t[k] > 0.0, k
    P = pg.PHASES
    assert t["p2p"] == 1e-3 * ph[P.index("p2p")]
    assert t["m2l"] == 1e-3 * ph[P.index("m2l")]
    gpu_task = 1e-3 * sum(ph[P.index(k)] for k in ("walk", "m2l", "l2l", "p2p"))
    assert gpu_task <= t["task"]
    gpu_prep = 1e-3 * sum(ph[P.index(k)] for k in ("upload", "tree", "upward"))
    assert gpu_prep <= t["prep"]
    assert abs(t["fmm"] - (t["prep"] + 1e-3 * ph[P.index("l2l")])) <= 1e-12 * t["fmm"]
